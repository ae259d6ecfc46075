/*
 * TaichiRTM oracle: a plain-C restatement of the reference's elastic P-SV/SH
 * staggered-grid time stepping, backward propagation and zero-lag imaging.
 *
 * TEST INFRASTRUCTURE.  Only the test suite, the smoke check and the CPU
 * baseline leg of the benchmark load this library, and only as the checker or
 * as the timed CPU stand-in for the reference; the product path never links it.
 *
 * It mirrors the reference's *unfused* per-step kernel sequence, one OpenMP
 * parallel loop per Taichi kernel, and is instantiated in float (the
 * reference's default_fp) and in double (the FP64 run of the same scheme used
 * to state rounding error).  Built with -ffp-contract=off so every float
 * expression is evaluated operand by operand in the order the reference
 * writes it; a second build with contraction on measures how much FMA
 * contraction (allowed by Taichi's fast_math default, rtm.py:108) can move
 * results.
 *
 * Parity pins (there are no reference tests or golden vectors): analytic 2-D
 * Green's functions for the SH and P-SV forward step, a reflection identity
 * tying the backward step to the forward one, closed forms for the sponge and
 * the harmonic/arithmetic averages, the image as an explicit sum over stored
 * wavefields, and linearity / time-shift invariants.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

enum { ST_U = 0, ST_V, ST_W, ST_SXX, ST_SZZ, ST_SXZ, ST_SYX, ST_SYZ, ST_COUNT };

typedef struct {
    int nx, nz, nt;
    double dx, dz, fs;
    int ns, nr;
    const int *src_loc; /* (ns, 2): column 0 = i (x), column 1 = j (z) */
    const int *rec_loc; /* (nr, 2) */
    int isnap;          /* forward snapshot interval */
    int sci;            /* stability_check_interval; <= 0 disables */
    const void *surface_matrix; /* NULL: flat free surface */
} oracle_problem;

typedef struct {
    const void *mu, *lam, *myx, *myz, *mxz, *inv_rho, *inv_rho_u, *inv_rho_w, *absorb;
} oracle_medium;

int oracle_abi_version(void) { return 1; }

#define REAL float
#define SFX _f32
#define EXPF expf
#define FABS fabsf
#include "fd_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXPF
#undef FABS

#define REAL double
#define SFX _f64
#define EXPF exp
#define FABS fabs
#include "fd_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXPF
#undef FABS
