/*
 * dual.c -- forward-mode (dual-number) derivatives of the oracle's energies.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker, never by the
 * product library.
 *
 * The reference's energies are plain Float64 Julia (lennard_jones and
 * vdw_energy, src/vdw_energetics.jl:19-47; _phi, src/electrostatics.jl:347-388;
 * Lorentz-Berthelot mixing, src/forcefield.jl:75-77).  Instantiated with dual
 * numbers (what ForwardDiff does to Julia code) they give exact derivatives,
 * with no step size.  This file is that instantiation in C: the arithmetic of
 * oracle.c's Widom energies and pose energies with every quantity a dual
 * number -- a value and ORC_DN derivative components -- and every operation
 * carrying its derivative rule (sum, product, quotient, sqrt, erfc, cos, sin).
 *
 *   orc_widom_param_grad: per insertion of the Widom stream, E_vdw, E_es and
 *     dE/d(eps_t, sigma_t, q_t) for listed framework species t: eps_t and
 *     sigma_t are the framework-side pure parameters fed through the mixing
 *     rule (the probe's held fixed), q_t a uniform shift of the charge of
 *     every framework atom of species t.  Optionally also the condition of
 *     each derivative as a sum: the sum of the magnitudes of its summands (per
 *     LJ pair, per real-space pair times the site charge, per site's Fourier
 *     part times the site charge), the scale of the rounding error of any
 *     evaluation that rounds each summand to a relative precision.
 *   orc_pose_grad: energies of explicit poses and dE/d(site xyz), Cartesian.
 *
 * Positions, cutoff tests and nearest images follow oracle.c exactly; the
 * Fourier part is the structure-factor form (equal to the reference's
 * e^{ik.r} form, pinned in tests/test_oracle_kat.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_cfg.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define R2_OVERLAP 0.1 /* src/vdw_energetics.jl:2 */
#define ORC_DN 24      /* derivative components of a dual number */

double orc_four_pi_eps0(void);
void orc_insertion_uniforms(uint64_t seed, uint64_t g, double u[6]);
void orc_arvo_rotation(const double u[3], double Rm[9]);

typedef struct {
    double v;
    double d[ORC_DN];
} dnum;

static inline dnum dn_c(double v) {
    dnum r;
    r.v = v;
    memset(r.d, 0, sizeof(r.d));
    return r;
}
static inline dnum dn_var(double v, int k) {
    dnum r = dn_c(v);
    r.d[k] = 1.0;
    return r;
}
static inline dnum dn_add(dnum a, dnum b) {
    dnum r;
    r.v = a.v + b.v;
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
static inline dnum dn_sub(dnum a, dnum b) {
    dnum r;
    r.v = a.v - b.v;
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
static inline dnum dn_mul(dnum a, dnum b) {
    dnum r;
    r.v = a.v * b.v;
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static inline dnum dn_div(dnum a, dnum b) {
    dnum r;
    r.v = a.v / b.v;
    const double ib2 = 1.0 / (b.v * b.v);
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) * ib2;
    return r;
}
static inline dnum dn_scale(dnum a, double s) {
    dnum r;
    r.v = a.v * s;
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = a.d[k] * s;
    return r;
}
static inline dnum dn_addc(dnum a, double c) {
    a.v += c;
    return a;
}
static inline dnum dn_chain(double v, double dv, dnum a) { /* f(a) with f(a.v) = v, f'(a.v) = dv */
    dnum r;
    r.v = v;
    for (int k = 0; k < ORC_DN; ++k) r.d[k] = dv * a.d[k];
    return r;
}
static inline dnum dn_sqrt(dnum a) {
    const double s = sqrt(a.v);
    return dn_chain(s, 0.5 / s, a);
}
static inline dnum dn_erfc(dnum a) {
    return dn_chain(erfc(a.v), -2.0 / sqrt(M_PI) * exp(-a.v * a.v), a);
}
static inline dnum dn_cos(dnum a) { return dn_chain(cos(a.v), -sin(a.v), a); }
static inline dnum dn_sin(dnum a) { return dn_chain(sin(a.v), cos(a.v), a); }

/* lennard_jones, src/vdw_energetics.jl:19-22 */
static dnum dn_lj(dnum r2, dnum sig2, dnum eps) {
    const dnum s = dn_div(sig2, r2);
    const dnum ratio = dn_mul(dn_mul(s, s), s);
    return dn_mul(dn_scale(eps, 4.0), dn_mul(ratio, dn_addc(ratio, -1.0)));
}

/* Xtals nearest_image! on the value; the shift is a constant of the pair */
static inline double image_shift(double d) {
    if (fabs(d) > 0.5) return (d > 0.0) ? 1.0 : -1.0;
    return 0.0;
}

static inline void mat3(const double m[9], const double v[3], double o[3]) {
    for (int i = 0; i < 3; ++i) o[i] = m[i] * v[0] + m[i + 3] * v[1] + m[i + 6] * v[2];
}

/* ------------------------------------------------------------------ */
/* Parameter derivatives of Widom energies                              */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t n_ft;                  /* framework species differentiated (<= ORC_DN / 3) */
    const int32_t* ft;             /* their force-field indices: components 3i + {0 eps, 1 sigma, 2 q} */
    const double* pure_eps;        /* per force-field species, K */
    const double* pure_sig;        /* per force-field species, A */
    int64_t n_unit;                /* unit-cell atoms (structure factors) */
    const double* unit_xf;         /* 3 x n_unit, unit-cell fractional */
    const double* unit_q;
    const int32_t* unit_species;
} orc_param_dual;

static int comp_of(const orc_param_dual* p, int32_t species) {
    for (int i = 0; i < p->n_ft; ++i)
        if (p->ft[i] == species) return i;
    return -1;
}

/* a framework charge as a dual number: q + (shift of its species' charges) */
static dnum dual_charge(const orc_param_dual* p, int32_t species, double q) {
    const int i = comp_of(p, species);
    return i >= 0 ? dn_var(q, 3 * i + 2) : dn_c(q);
}

/* The Lorentz-Berthelot pair of framework species t (dual pure parameters)
 * and probe species s (constants): eps_ts = sqrt(eps_t eps_s),
 * sigma_ts^2 = ((sigma_t + sigma_s) / 2)^2. */
static void dual_mix(const orc_param_dual* p, int32_t t, int32_t s, dnum* eps, dnum* sig2) {
    const int i = comp_of(p, t);
    const dnum et = i >= 0 ? dn_var(p->pure_eps[t], 3 * i) : dn_c(p->pure_eps[t]);
    const dnum st = i >= 0 ? dn_var(p->pure_sig[t], 3 * i + 1) : dn_c(p->pure_sig[t]);
    *eps = dn_sqrt(dn_scale(et, p->pure_eps[s]));
    const dnum sts = dn_scale(dn_addc(st, p->pure_sig[s]), 0.5);
    *sig2 = dn_mul(sts, sts);
}

int32_t orc_widom_param_grad(const orc_widom_cfg* c, const orc_param_dual* p, uint64_t g0, int64_t count,
                             double* out_vdw, double* out_es, double* out_d, double* out_cond) {
    if (p->n_ft * 3 > ORC_DN || c->n_lj > 16 || c->n_q > 16) return 1;
    const int n_sp = c->n_species;
    /* dual mixed LJ tables [framework species][probe species] */
    dnum* eps_t = (dnum*)malloc(sizeof(dnum) * (size_t)n_sp * n_sp);
    dnum* sig2_t = (dnum*)malloc(sizeof(dnum) * (size_t)n_sp * n_sp);
    for (int t = 0; t < n_sp; ++t)
        for (int s = 0; s < n_sp; ++s) dual_mix(p, t, s, &eps_t[t * n_sp + s], &sig2_t[t * n_sp + s]);
    /* dual structure factors on the kept unit-cell k-vectors (orc_sk_build's
     * folding of the replicated-box list), every unit atom included */
    const int64_t nr = (int64_t)c->rep[0] * c->rep[1] * c->rep[2];
    int32_t* m = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(c->nk + 1));
    dnum* are = (dnum*)malloc(sizeof(dnum) * (size_t)(c->nk + 1));
    dnum* aim = (dnum*)malloc(sizeof(dnum) * (size_t)(c->nk + 1));
    int64_t nks = 0;
    if (c->charged) {
        for (int64_t k = 0; k < c->nk; ++k) {
            const int na = c->kidx[3 * k], nb = c->kidx[3 * k + 1], nc = c->kidx[3 * k + 2];
            if (na % c->rep[0] || nb % c->rep[1] || nc % c->rep[2]) continue;
            const int ma = na / c->rep[0], mb = nb / c->rep[1], mc = nc / c->rep[2];
            dnum sre = dn_c(0.0), sim = dn_c(0.0);
            for (int64_t u = 0; u < p->n_unit; ++u) {
                double t = ma * p->unit_xf[3 * u] + mb * p->unit_xf[3 * u + 1] + mc * p->unit_xf[3 * u + 2];
                t -= nearbyint(t);
                const double ang = 2.0 * M_PI * t;
                const dnum q = dual_charge(p, p->unit_species[u], p->unit_q[u]);
                sre = dn_add(sre, dn_scale(q, cos(ang)));
                sim = dn_add(sim, dn_scale(q, sin(ang)));
            }
            m[3 * nks] = ma;
            m[3 * nks + 1] = mb;
            m[3 * nks + 2] = mc;
            are[nks] = dn_scale(sre, c->wt[k] * (double)nr);
            aim[nks] = dn_scale(sim, c->wt[k] * (double)nr);
            ++nks;
        }
    }
    const double fpe = orc_four_pi_eps0();
    const int nd = 3 * p->n_ft;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ii = 0; ii < count; ++ii) {
        const uint64_t g = g0 + (uint64_t)ii;
        const int rotate = (c->n_lj + c->n_q) > 1;
        double u[6], Rm[9];
        orc_insertion_uniforms(c->seed, g, u);
        if (rotate) orc_arvo_rotation(&u[3], Rm);
        double sxf[3 * 16], qxf[3 * 16];
        for (int pass = 0; pass < 2; ++pass) {
            const int ns = pass ? c->n_q : c->n_lj;
            const double* offs = pass ? c->q_off : c->lj_off;
            double* dst = pass ? qxf : sxf;
            for (int s = 0; s < ns; ++s) {
                double y[3], f[3];
                if (rotate) mat3(Rm, &offs[3 * s], y);
                else { y[0] = offs[3 * s]; y[1] = offs[3 * s + 1]; y[2] = offs[3 * s + 2]; }
                mat3(c->c_to_f, y, f);
                for (int d = 0; d < 3; ++d) dst[3 * s + d] = u[d] + f[d];
            }
        }
        dnum vdw = dn_c(0.0);
        double cond[ORC_DN] = {0.0};  /* sum of |summand derivative| */
        int overlap = 0;
        for (int s = 0; s < c->n_lj && !overlap; ++s) {
            const int ps = c->lj_species[s];
            for (int64_t j = 0; j < c->n_rep; ++j) {
                double dl[3], cc[3];
                for (int k = 0; k < 3; ++k) {
                    const double d = c->xf[3 * j + k] - sxf[3 * s + k];
                    dl[k] = d - image_shift(d);
                }
                mat3(c->F, dl, cc);
                const double r2 = cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2];
                if (r2 < c->r2_cutoff) {
                    if (r2 < R2_OVERLAP) { overlap = 1; break; }
                    const int t = c->species[j];
                    const dnum e = dn_lj(dn_c(r2), sig2_t[t * n_sp + ps], eps_t[t * n_sp + ps]);
                    vdw = dn_add(vdw, e);
                    for (int k = 0; k < nd; ++k) cond[k] += fabs(e.d[k]);
                }
            }
        }
        dnum es = dn_c(0.0);
        if (c->charged && !overlap) {
            for (int s = 0; s < c->n_q; ++s) {
                dnum sr = dn_c(0.0);
                for (int64_t j = 0; j < c->n_rep; ++j) {
                    double dl[3], cc[3];
                    for (int k = 0; k < 3; ++k) {
                        const double d = c->xf[3 * j + k] - qxf[3 * s + k];
                        dl[k] = d - image_shift(d);
                    }
                    mat3(c->F, dl, cc);
                    const double r = sqrt(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
                    if (r < c->sr_cutoff) {
                        const dnum q = dual_charge(p, c->species[j], c->q[j]);
                        const dnum e = dn_scale(q, erfc(r * c->alpha) / r / fpe);
                        sr = dn_add(sr, e);
                        for (int k = 0; k < nd; ++k) cond[k] += fabs(c->q_site[s] * e.d[k]);
                    }
                }
                const double xu[3] = {qxf[3 * s] * c->rep[0], qxf[3 * s + 1] * c->rep[1], qxf[3 * s + 2] * c->rep[2]};
                dnum lr = dn_c(0.0);
                for (int64_t k = 0; k < nks; ++k) {
                    double t = m[3 * k] * xu[0] + m[3 * k + 1] * xu[1] + m[3 * k + 2] * xu[2];
                    t -= nearbyint(t);
                    const double ang = 2.0 * M_PI * t;
                    lr = dn_add(lr, dn_add(dn_scale(are[k], cos(ang)), dn_scale(aim[k], sin(ang))));
                }
                for (int k = 0; k < nd; ++k) cond[k] += fabs(c->q_site[s] * lr.d[k]);
                es = dn_add(es, dn_scale(dn_add(sr, lr), c->q_site[s]));
            }
        }
        out_vdw[ii] = overlap ? INFINITY : vdw.v;
        out_es[ii] = overlap ? 0.0 : es.v;
        for (int k = 0; k < nd; ++k) out_d[ii * nd + k] = overlap ? 0.0 : vdw.d[k] + es.d[k];
        if (out_cond)
            for (int k = 0; k < nd; ++k) out_cond[ii * nd + k] = overlap ? 0.0 : cond[k];
    }
    free(eps_t);
    free(sig2_t);
    free(m);
    free(are);
    free(aim);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Position derivatives: explicit poses                                 */
/* ------------------------------------------------------------------ */
/* n_sites sites per pose, each with an LJ species (-1: none) and a charge;
 * site_xf: n_poses x n_sites x 3 fractional in the replicated box.  out_e:
 * energy per pose (K; +Inf on overlap); out_g: n_poses x n_sites x 3,
 * dE/d(site Cartesian xyz) (K/A; 0 on overlap).  LJ from the cfg's mixed
 * tables; real-space Ewald over the nonzero framework charges; Fourier in
 * the structure-factor form (a_re, a_im on m_sk). */
int32_t orc_pose_grad(const orc_widom_cfg* c, int32_t n_sites, const int32_t* site_species, const double* site_q,
                      int64_t n_poses, const double* site_xf, double* out_e, double* out_g) {
    if (3 * n_sites > ORC_DN) return 1;
    const double fpe = orc_four_pi_eps0();
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pi = 0; pi < n_poses; ++pi) {
        dnum xf[8][3];
        for (int s = 0; s < n_sites; ++s) {
            double x[3];
            mat3(c->F, &site_xf[(pi * n_sites + s) * 3], x);
            dnum xc[3];
            for (int d = 0; d < 3; ++d) xc[d] = dn_var(x[d], 3 * s + d);
            for (int d = 0; d < 3; ++d)
                xf[s][d] = dn_add(dn_add(dn_scale(xc[0], c->c_to_f[d]), dn_scale(xc[1], c->c_to_f[d + 3])),
                                  dn_scale(xc[2], c->c_to_f[d + 6]));
        }
        dnum e = dn_c(0.0);
        int overlap = 0;
        for (int s = 0; s < n_sites && !overlap; ++s) {
            const int ps = site_species[s];
            if (ps >= 0) {
                for (int64_t j = 0; j < c->n_rep; ++j) {
                    dnum dl[3];
                    for (int k = 0; k < 3; ++k) {
                        const dnum d = dn_sub(dn_c(c->xf[3 * j + k]), xf[s][k]);
                        dl[k] = dn_addc(d, -image_shift(d.v));
                    }
                    dnum cc[3];
                    for (int i = 0; i < 3; ++i)
                        cc[i] = dn_add(dn_add(dn_scale(dl[0], c->F[i]), dn_scale(dl[1], c->F[i + 3])),
                                       dn_scale(dl[2], c->F[i + 6]));
                    const dnum r2 = dn_add(dn_add(dn_mul(cc[0], cc[0]), dn_mul(cc[1], cc[1])), dn_mul(cc[2], cc[2]));
                    if (r2.v < c->r2_cutoff) {
                        if (r2.v < R2_OVERLAP) { overlap = 1; break; }
                        const int64_t pp = (int64_t)ps * c->n_species + c->species[j];
                        e = dn_add(e, dn_lj(r2, dn_c(c->sig2[pp]), dn_c(c->eps[pp])));
                    }
                }
            }
            if (overlap || site_q[s] == 0.0 || !c->charged) continue;
            dnum sr = dn_c(0.0);
            for (int64_t j = 0; j < c->nq; ++j) {
                dnum dl[3];
                for (int k = 0; k < 3; ++k) {
                    const dnum d = dn_sub(dn_c(c->xq[3 * j + k]), xf[s][k]);
                    dl[k] = dn_addc(d, -image_shift(d.v));
                }
                dnum cc[3];
                for (int i = 0; i < 3; ++i)
                    cc[i] = dn_add(dn_add(dn_scale(dl[0], c->F[i]), dn_scale(dl[1], c->F[i + 3])),
                                   dn_scale(dl[2], c->F[i + 6]));
                const dnum r = dn_sqrt(dn_add(dn_add(dn_mul(cc[0], cc[0]), dn_mul(cc[1], cc[1])), dn_mul(cc[2], cc[2])));
                if (r.v < c->sr_cutoff)
                    sr = dn_add(sr, dn_scale(dn_div(dn_erfc(dn_scale(r, c->alpha)), r), c->qq[j] / fpe));
            }
            dnum lr = dn_c(0.0);
            for (int64_t k = 0; k < c->nk_sk; ++k) {
                dnum t = dn_add(dn_add(dn_scale(xf[s][0], (double)c->m_sk[3 * k] * c->rep[0]),
                                       dn_scale(xf[s][1], (double)c->m_sk[3 * k + 1] * c->rep[1])),
                                dn_scale(xf[s][2], (double)c->m_sk[3 * k + 2] * c->rep[2]));
                t = dn_addc(t, -nearbyint(t.v));
                const dnum ang = dn_scale(t, 2.0 * M_PI);
                lr = dn_add(lr, dn_add(dn_scale(dn_cos(ang), c->a_re[k]), dn_scale(dn_sin(ang), c->a_im[k])));
            }
            e = dn_add(e, dn_scale(dn_add(sr, lr), site_q[s]));
        }
        out_e[pi] = overlap ? INFINITY : e.v;
        for (int k = 0; k < 3 * n_sites; ++k) out_g[pi * 3 * n_sites + k] = overlap ? 0.0 : e.d[k];
    }
    return 0;
}
