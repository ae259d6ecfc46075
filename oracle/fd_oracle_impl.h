/*
 * Body of the CPU oracle, instantiated twice by fd_oracle.c (REAL = float with
 * suffix _f32, REAL = double with suffix _f64).  TEST INFRASTRUCTURE ONLY.
 *
 * Each function restates one Taichi kernel (or the time loop that sequences
 * them) of the reference, unfused and in the reference's own order, with the
 * reference's floating-point expression order kept operand for operand.
 * Citations are to mkt-kuno/TaichiRTM src/.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define AT(f, i, j) ((REAL *)(f))[(size_t)(i) * (size_t)nz + (size_t)(j)]

/* forward_modeling.py:203-241 (_init_absorbing_kernel / _init_absorbing).
 * Three sequential offloaded loops, left -> right -> bottom; later writes win. */
void FN(oracle_sponge)(int nx, int nz, int fw, double a_in, void *coeff)
{
    REAL a = (REAL)a_in; /* a: ti.f32 kernel argument */
    REAL a_sq = a * a;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) AT(coeff, i, j) = (REAL)1.0;
    for (int i = 0; i < fw; i++) /* left */
        for (int j = 0; j < nz; j++)
            if (j < nz - i - 1) {
                int diff = fw - i;
                AT(coeff, i, j) = EXPF((-a_sq * (REAL)diff) * (REAL)diff);
            }
    for (int i = 0; i < fw; i++) /* right */
        for (int j = 0; j < nz; j++) {
            int ii = nx - i - 1;
            if (j < nz - i - 1) {
                int diff = fw - i;
                AT(coeff, ii, j) = EXPF((-a_sq * (REAL)diff) * (REAL)diff);
            }
        }
    for (int i = 0; i < nx; i++) /* bottom */
        for (int j = 0; j < fw; j++) {
            int jj = nz - j - 1;
            if (i >= j && i < nx - j) {
                int diff = fw - j;
                AT(coeff, i, jj) = EXPF((-a_sq * (REAL)diff) * (REAL)diff);
            }
        }
}

/* forward_modeling.py:243-307 (_compute_shear_avg, _compute_rho_avg,
 * _compute_inv_rho).  Averages on the interior only (borders keep 0); the
 * reciprocals on the full grid (so 1/0 = +inf on the averaged borders). */
void FN(oracle_prepare)(int nx, int nz, const void *mu, const void *rho, void *myx, void *myz,
                        void *mxz, void *rho_u, void *rho_w, void *inv_rho, void *inv_rho_u,
                        void *inv_rho_w)
{
    const REAL two = (REAL)2.0, four = (REAL)4.0, half = (REAL)0.5;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            AT(myx, i, j) = 0; AT(myz, i, j) = 0; AT(mxz, i, j) = 0;
            AT(rho_u, i, j) = 0; AT(rho_w, i, j) = 0;
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
            REAL a = AT(mu, i, j), b = AT(mu, i + 1, j), c = AT(mu, i, j + 1), d = AT(mu, i + 1, j + 1);
            AT(myx, i, j) = two * a * b / (a + b);
            AT(myz, i, j) = two * a * c / (a + c);
            REAL product = a * b * c * d;
            REAL sum_inv_prod = b * c * d + a * c * d + a * b * d + a * b * c;
            AT(mxz, i, j) = four * product / sum_inv_prod;
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
            REAL r = AT(rho, i, j), rx = AT(rho, i + 1, j), rz = AT(rho, i, j + 1);
            AT(rho_u, i, j) = half * (r + rx);
            AT(rho_w, i, j) = half * (r + rz);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            AT(inv_rho, i, j) = (REAL)1.0 / AT(rho, i, j);
            AT(inv_rho_u, i, j) = (REAL)1.0 / AT(rho_u, i, j);
            AT(inv_rho_w, i, j) = (REAL)1.0 / AT(rho_w, i, j);
        }
}

/* forward_modeling.py:404-419 (_set_free_surface / _set_surface_boundary) */
static void FN(bc)(int nx, int nz, void **st, const void *sm)
{
    if (sm) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < nz; j++) {
                REAL s = AT(sm, i, j);
                AT(st[ST_SYZ], i, j) *= s;
                AT(st[ST_SXZ], i, j) *= s;
                AT(st[ST_SZZ], i, j) *= s;
#if ORACLE_PLANT == 7
                AT(st[ST_SXX], i, j) *= s;
#endif
            }
    } else {
        for (int i = 0; i < nx; i++) {
            AT(st[ST_SYZ], i, 0) = 0;
            AT(st[ST_SXZ], i, 0) = 0;
            AT(st[ST_SZZ], i, 0) = 0;
        }
    }
}

/* forward_modeling.py:386-402 / backward_modeling.py:382-398 (_apply_absorbing) */
static void FN(sponge_apply)(int nx, int nz, void **st, const void *coeff)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            REAL c = AT(coeff, i, j);
            AT(st[ST_U], i, j) *= c;
            AT(st[ST_V], i, j) *= c;
            AT(st[ST_W], i, j) *= c;
            AT(st[ST_SXX], i, j) *= c;
            AT(st[ST_SXZ], i, j) *= c;
            AT(st[ST_SZZ], i, j) *= c;
            AT(st[ST_SYX], i, j) *= c;
            AT(st[ST_SYZ], i, j) *= c;
        }
}

/* forward_modeling.py:455-488 / backward_modeling.py:417-450 (_check_finite):
 * NaN or |x| > 1e30 on u, v, w; highest code wins. */
static int FN(check_finite)(int nx, int nz, void **st, int base)
{
    int result = 0;
    const REAL big = (REAL)1e30;
#pragma omp parallel for schedule(static) reduction(max : result)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            REAL u = AT(st[ST_U], i, j), v = AT(st[ST_V], i, j), w = AT(st[ST_W], i, j);
            if (u != u || FABS(u) > big) result = result > base + 1 ? result : base + 1;
            if (v != v || FABS(v) > big) result = result > base + 2 ? result : base + 2;
            if (w != w || FABS(w) > big) result = result > base + 3 ? result : base + 3;
        }
    return result;
}

/* forward_modeling.py:309-347 (_update_velocity) */
static void FN(velocity_fw)(int nx, int nz, void **st, const oracle_medium *m, REAL dt, REAL idx, REAL idz)
{
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
            REAL sxx_x = (AT(st[ST_SXX], i, j) - AT(st[ST_SXX], i - 1, j)) * idx;
            REAL szz_z = (AT(st[ST_SZZ], i, j) - AT(st[ST_SZZ], i, j - 1)) * idz;
            REAL sxz_x = (AT(st[ST_SXZ], i, j) - AT(st[ST_SXZ], i - 1, j)) * idx;
            REAL sxz_z = (AT(st[ST_SXZ], i, j) - AT(st[ST_SXZ], i, j - 1)) * idz;
            AT(st[ST_U], i, j) += (sxx_x + sxz_z) * dt * AT(m->inv_rho_u, i, j);
            AT(st[ST_W], i, j) += (sxz_x + szz_z) * dt * AT(m->inv_rho_w, i, j);
            REAL syx_x = (AT(st[ST_SYX], i, j) - AT(st[ST_SYX], i - 1, j)) * idx;
            REAL syz_z = (AT(st[ST_SYZ], i, j) - AT(st[ST_SYZ], i, j - 1)) * idz;
            AT(st[ST_V], i, j) += (syx_x + syz_z) * dt * AT(m->inv_rho, i, j);
        }
}

/* forward_modeling.py:349-384 (_update_stress) */
static void FN(stress_fw)(int nx, int nz, void **st, const oracle_medium *m, REAL dt, REAL idx, REAL idz)
{
    const REAL two = (REAL)2.0;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
            REAL u_x = (AT(st[ST_U], i + 1, j) - AT(st[ST_U], i, j)) * idx;
#if ORACLE_PLANT == 3
            REAL u_z = (AT(st[ST_U], i + 1, j) - AT(st[ST_U], i, j)) * idz;
#else
            REAL u_z = (AT(st[ST_U], i, j + 1) - AT(st[ST_U], i, j)) * idz;
#endif
            REAL w_x = (AT(st[ST_W], i + 1, j) - AT(st[ST_W], i, j)) * idx;
            REAL w_z = (AT(st[ST_W], i, j + 1) - AT(st[ST_W], i, j)) * idz;
            REAL div_uw = u_x + w_z;
            REAL lam = AT(m->lam, i, j), mu = AT(m->mu, i, j);
            AT(st[ST_SXX], i, j) += dt * (lam * div_uw + two * mu * u_x);
            AT(st[ST_SZZ], i, j) += dt * (lam * div_uw + two * mu * w_z);
#if ORACLE_PLANT == 1
            AT(st[ST_SXZ], i, j) -= dt * AT(m->mxz, i, j) * (u_z + w_x);
#else
            AT(st[ST_SXZ], i, j) += dt * AT(m->mxz, i, j) * (u_z + w_x);
#endif
            REAL v_x = (AT(st[ST_V], i + 1, j) - AT(st[ST_V], i, j)) * idx;
            REAL v_z = (AT(st[ST_V], i, j + 1) - AT(st[ST_V], i, j)) * idz;
            AT(st[ST_SYX], i, j) += dt * AT(m->myx, i, j) * v_x;
            AT(st[ST_SYZ], i, j) += dt * AT(m->myz, i, j) * v_z;
        }
}

/* backward_modeling.py:312-345 (_update_velocity_backward) */
static void FN(velocity_bw)(int nx, int nz, void **st, const oracle_medium *m, REAL dt, REAL idx, REAL idz)
{
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
#if ORACLE_PLANT == 4
            REAL sxx_x = (AT(st[ST_SXX], i, j) - AT(st[ST_SXX], i - 1, j)) * idx;
#else
            REAL sxx_x = (AT(st[ST_SXX], i + 1, j) - AT(st[ST_SXX], i, j)) * idx;
#endif
            REAL szz_z = (AT(st[ST_SZZ], i, j + 1) - AT(st[ST_SZZ], i, j)) * idz;
            REAL sxz_x = (AT(st[ST_SXZ], i + 1, j) - AT(st[ST_SXZ], i, j)) * idx;
            REAL sxz_z = (AT(st[ST_SXZ], i, j + 1) - AT(st[ST_SXZ], i, j)) * idz;
            AT(st[ST_U], i, j) += -(sxx_x + sxz_z) * dt * AT(m->inv_rho_u, i, j);
            AT(st[ST_W], i, j) += -(sxz_x + szz_z) * dt * AT(m->inv_rho_w, i, j);
            REAL syx_x = (AT(st[ST_SYX], i + 1, j) - AT(st[ST_SYX], i, j)) * idx;
            REAL syz_z = (AT(st[ST_SYZ], i, j + 1) - AT(st[ST_SYZ], i, j)) * idz;
            AT(st[ST_V], i, j) += -(syx_x + syz_z) * dt * AT(m->inv_rho, i, j);
        }
}

/* backward_modeling.py:347-380 (_update_stress_backward) */
static void FN(stress_bw)(int nx, int nz, void **st, const oracle_medium *m, REAL dt, REAL idx, REAL idz)
{
    const REAL two = (REAL)2.0;
    const REAL ndt = -dt;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < nz - 1; j++) {
            REAL u_x = (AT(st[ST_U], i, j) - AT(st[ST_U], i - 1, j)) * idx;
            REAL u_z = (AT(st[ST_U], i, j) - AT(st[ST_U], i, j - 1)) * idz;
            REAL w_x = (AT(st[ST_W], i, j) - AT(st[ST_W], i - 1, j)) * idx;
            REAL w_z = (AT(st[ST_W], i, j) - AT(st[ST_W], i, j - 1)) * idz;
            REAL div_uw = u_x + w_z;
            REAL lam = AT(m->lam, i, j), mu = AT(m->mu, i, j);
            AT(st[ST_SXX], i, j) += ndt * (lam * div_uw + two * mu * u_x);
            AT(st[ST_SZZ], i, j) += ndt * (lam * div_uw + two * mu * w_z);
            AT(st[ST_SXZ], i, j) += ndt * AT(m->mxz, i, j) * (u_z + w_x);
            REAL v_x = (AT(st[ST_V], i, j) - AT(st[ST_V], i - 1, j)) * idx;
            REAL v_z = (AT(st[ST_V], i, j) - AT(st[ST_V], i, j - 1)) * idz;
            AT(st[ST_SYX], i, j) += ndt * AT(m->myx, i, j) * v_x;
            AT(st[ST_SYZ], i, j) += ndt * AT(m->myz, i, j) * v_z;
        }
}

/* forward_modeling.py:504-575 (ForwardModeling.run), steps [it0, it1).
 * Per step: BC -> V -> source(it) -> S -> sponge -> record(it)
 *           -> finite check if it % sci == 0 (return code) -> snapshot.
 * Returns 0 or the finite-check code (1..3); *it_stop gets the step index. */
int FN(oracle_forward)(const oracle_problem *p, void **st, const oracle_medium *m,
                       const void *wav_u, const void *wav_v, const void *wav_w,
                       void *seis_u, void *seis_v, void *seis_w,
                       int save, void *snap_u, void *snap_v, void *snap_w, int *isnaps,
                       int it0, int it1, int *it_stop)
{
    const int nx = p->nx, nz = p->nz, nt = p->nt;
    const double dt_d = 1.0 / p->fs;
    const REAL dt = (REAL)dt_d, idx = (REAL)(1.0 / p->dx), idz = (REAL)(1.0 / p->dz);
#if ORACLE_PLANT == 2
    const REAL factor = (REAL)(dt_d / (p->dx * p->dz));
#else
    const REAL factor = (REAL)(dt_d * p->dx * p->dz);
#endif
    const int num_snaps = p->isnap > 0 ? nt / p->isnap : 0;
    for (int it = it0; it < it1; it++) {
#if ORACLE_PLANT == 8
        FN(velocity_fw)(nx, nz, st, m, dt, idx, idz);
        FN(bc)(nx, nz, st, p->surface_matrix);
#else
        FN(bc)(nx, nz, st, p->surface_matrix);
        FN(velocity_fw)(nx, nz, st, m, dt, idx, idz);
#endif
        for (int k = 0; k < p->ns; k++) { /* forward_modeling.py:421-443 */
            int i = p->src_loc[2 * k], j = p->src_loc[2 * k + 1];
            REAL iru = AT(m->inv_rho_u, i, j), ir = AT(m->inv_rho, i, j), irw = AT(m->inv_rho_w, i, j);
            AT(st[ST_U], i, j) += ((const REAL *)wav_u)[(size_t)k * nt + it] * factor * iru;
            AT(st[ST_V], i, j) += ((const REAL *)wav_v)[(size_t)k * nt + it] * factor * ir;
            AT(st[ST_W], i, j) += ((const REAL *)wav_w)[(size_t)k * nt + it] * factor * irw;
        }
        FN(stress_fw)(nx, nz, st, m, dt, idx, idz);
        FN(sponge_apply)(nx, nz, st, m->absorb);
        for (int l = 0; l < p->nr; l++) { /* forward_modeling.py:445-453 */
            int i = p->rec_loc[2 * l], j = p->rec_loc[2 * l + 1];
            ((REAL *)seis_u)[(size_t)l * nt + it] = AT(st[ST_U], i, j);
            ((REAL *)seis_v)[(size_t)l * nt + it] = AT(st[ST_V], i, j);
            ((REAL *)seis_w)[(size_t)l * nt + it] = AT(st[ST_W], i, j);
        }
        if (p->sci > 0 && it % p->sci == 0) {
            int flag = FN(check_finite)(nx, nz, st, 0);
            if (flag) { if (it_stop) *it_stop = it; return flag; }
        }
        if (save && it % p->isnap == 0 && it != 0) { /* forward_modeling.py:568-572 */
            int s = it / p->isnap - 1;
            if (s < num_snaps) {
                for (int i = 0; i < nx; i++)
                    for (int j = 0; j < nz; j++) {
                        size_t o = ((size_t)i * nz + j) * (size_t)num_snaps + s;
                        ((REAL *)snap_u)[o] = AT(st[ST_U], i, j);
                        ((REAL *)snap_v)[o] = AT(st[ST_V], i, j);
                        ((REAL *)snap_w)[o] = AT(st[ST_W], i, j);
                    }
                isnaps[s] = it;
            }
        }
    }
    if (it_stop) *it_stop = it1;
    return 0;
}

/* backward_modeling.py:498-579 (BackwardModeling.run_calc), steps [it0, it1).
 * Per step, t = nt-1-it: BC -> V_b -> inject obs(t) -> S_b -> sponge -> synsrc(t)
 *   -> image (t>0, t%isnap==0, valid snapshot, isnaps[s]==t) -> finite check. */
int FN(oracle_backward)(const oracle_problem *p, void **st, const oracle_medium *m,
                        const void *obs_u, const void *obs_v, const void *obs_w,
                        void *syn_u, void *syn_v, void *syn_w,
                        const void *snap_u, const void *snap_v, const void *snap_w,
                        const int *isnaps, int num_snaps, int isnap_interval,
                        void *res_u, void *res_v, void *res_w, int it0, int it1, int *it_stop)
{
    const int nx = p->nx, nz = p->nz, nt = p->nt;
    const double dt_d = 1.0 / p->fs;
    const REAL dt = (REAL)dt_d, idx = (REAL)(1.0 / p->dx), idz = (REAL)(1.0 / p->dz);
    for (int it = it0; it < it1; it++) {
        int t = (nt - 1) - it;
        FN(bc)(nx, nz, st, p->surface_matrix);
        FN(velocity_bw)(nx, nz, st, m, dt, idx, idz);
        for (int k = 0; k < p->nr; k++) { /* backward_modeling.py:452-460 */
            int i = p->rec_loc[2 * k], j = p->rec_loc[2 * k + 1];
#if ORACLE_PLANT == 5
            if (t == 0) continue;
            t -= 1;
#endif
            AT(st[ST_U], i, j) += ((const REAL *)obs_u)[(size_t)k * nt + t];
            AT(st[ST_V], i, j) += ((const REAL *)obs_v)[(size_t)k * nt + t];
            AT(st[ST_W], i, j) += ((const REAL *)obs_w)[(size_t)k * nt + t];
#if ORACLE_PLANT == 5
            t += 1;
#endif
        }
        FN(stress_bw)(nx, nz, st, m, dt, idx, idz);
        FN(sponge_apply)(nx, nz, st, m->absorb);
        if (syn_u) {
            for (int l = 0; l < p->ns; l++) { /* backward_modeling.py:462-470 */
                int i = p->src_loc[2 * l], j = p->src_loc[2 * l + 1];
                ((REAL *)syn_u)[(size_t)l * nt + t] = AT(st[ST_U], i, j);
                ((REAL *)syn_v)[(size_t)l * nt + t] = AT(st[ST_V], i, j);
                ((REAL *)syn_w)[(size_t)l * nt + t] = AT(st[ST_W], i, j);
            }
        }
        if (t > 0 && isnap_interval > 0 && t % isnap_interval == 0) { /* :556-564 */
            int s = t / isnap_interval - 1;
#if ORACLE_PLANT == 6
            if (s >= 1 && s < num_snaps && isnaps[s] == t) {
                s -= 1;
#else
            if (s >= 0 && s < num_snaps && isnaps[s] == t) {
#endif
#pragma omp parallel for schedule(static)
                for (int i = 0; i < nx; i++)
                    for (int j = 0; j < nz; j++) { /* :481-488 */
                        size_t o = ((size_t)i * nz + j) * (size_t)num_snaps + s;
                        AT(res_u, i, j) += ((const REAL *)snap_u)[o] * AT(st[ST_U], i, j);
                        AT(res_v, i, j) += ((const REAL *)snap_v)[o] * AT(st[ST_V], i, j);
                        AT(res_w, i, j) += ((const REAL *)snap_w)[o] * AT(st[ST_W], i, j);
                    }
            }
        }
        if (p->sci > 0 && it % p->sci == 0) {
            int flag = FN(check_finite)(nx, nz, st, 3);
            if (flag) { if (it_stop) *it_stop = it; return flag; }
        }
    }
    if (it_stop) *it_stop = it1;
    return 0;
}

#undef AT
#undef FN
#undef CAT
#undef CAT2
