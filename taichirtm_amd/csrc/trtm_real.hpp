// Kernel argument blocks over a real type REAL (float: product path; double:
// the FP64 runs that check the FWI gradient).
#pragma once

#include "trtm_common.hpp"

namespace trtm {

// The same kernels serve float (the product path) and double (the FP64 runs
// that check the FWI gradient); KA/IA/SIO are KArgs/InjArgs/StepIO over REAL.
template <class R>
struct KA {
    int nx, nz, nt, fw;
    R dt, idx, idz, factor;
    const R* tab;
    const R *lam, *mu, *mxz, *myx, *myz, *ir, *iru, *irw;
    const R* sm;
    R *u, *v, *w, *sxx, *szz, *sxz, *syx, *syz;
};

template <class R>
struct IA {
    int n_groups, col;
    const int *loc, *order, *start;
    const R *su, *sv, *sw;
};

template <class R>
struct SIO {
    int snap_slot, image_slot, check_slot, check_base, rec_row;
    R *snap_u, *snap_v, *snap_w, *image_u, *image_v, *image_w;
    int* check_flags;
    int n_rec;
    const int* rec_loc;
    R *rec_u, *rec_v, *rec_w;
};

template <class R>
__device__ __forceinline__ R sponge_at_t(int i, int j, int nx, int nz, int fw,
                                         const R* __restrict__ tab) {
    const int jb = nz - 1 - j;
    if (jb < fw && i >= jb && i < nx - jb) return tab[fw - jb];
    const int ir = nx - 1 - i;
    if (ir < fw && j < nz - ir - 1) return tab[fw - ir];
    if (i < fw && j < nz - i - 1) return tab[fw - i];
    return R(1);
}

template <class R>
__device__ __forceinline__ bool bad_value(R x) {
    return (x != x) || (fabs((double)x) > (double)1e30f);
}

}  // namespace trtm
