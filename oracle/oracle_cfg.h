/*
 * oracle_cfg.h -- the whole-system configuration shared by the oracle's
 * Widom / grid functions (oracle.c) and its dual-number derivatives
 * (dual.c).  TEST INFRASTRUCTURE ONLY.
 *
 * Framework: replicated atoms (fractional coordinates of the replicated
 * box), their species and charges; the nonzero charges again as (xq, qq).
 * Probe: LJ sites and charge sites as Cartesian offsets from the centre of
 * mass.  es_mode 0: structure-factor Fourier form (a_* from orc_sk_build);
 * es_mode 1: the reference's e^{ik.r} form (kidx/wt on the replicated box).
 */
#ifndef ORACLE_CFG_H
#define ORACLE_CFG_H

#include <stdint.h>

typedef struct {
    int64_t n_rep;
    const double* xf; const int32_t* species; const double* q;
    int64_t nq; const double* xq; const double* qq; /* nonzero charges only */
    double F[9], c_to_f[9];
    int32_t rep[3];
    int32_t n_species; const double* eps; const double* sig2; double r2_cutoff;
    int32_t n_lj; const double* lj_off; const int32_t* lj_species;
    int32_t n_q; const double* q_off; const double* q_site;
    int32_t charged; double alpha; double sr_cutoff;
    int32_t es_mode;
    int64_t nk_sk; const int32_t* m_sk; const double* a_re; const double* a_im;
    int32_t kreps[3]; int64_t nk; const int32_t* kidx; const double* wt;
    uint64_t seed;
} orc_widom_cfg;

#endif
