/*
 * gcmc.c -- CPU FP64 restatement of the reference's grand-canonical Monte
 * Carlo chain (TEST INFRASTRUCTURE ONLY: the checker of libpmhip.so's GCMC
 * chains, loaded by tests/ through oracle/gcmc.py).
 *
 * One chain, the reference's move set and acceptance rules
 * (src/gcmc.jl:175-817), its energies (src/gcmc.jl:111-124,
 * src/vdw_energetics.jl:19-47,85-120) and helpers (src/mc_helpers.jl:17-190,
 * src/molecule.jl:233-275), driven by the counter-based random stream the
 * device uses (Philox4x32-10; move t of chain c reads 10 uniforms), so that a
 * device chain and this chain make the same proposals and, up to FP64
 * rounding in the energies, the same decisions.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define GC_NS 4
#define GC_MS 8
#define GC_NB 5
#define GC_NM 6
#define R2_OVERLAP 0.1
#define KB 1.38064852e7 /* Pa A^3 / K, src/PorousMaterials.jl:14 */
#define GCMC_TAG 0x47434D00u

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]);
void orc_arvo_rotation(const double u[3], double Rm[9]);

enum { MV_INSERTION = 0, MV_DELETION = 1, MV_TRANSLATION = 2, MV_ROTATION = 3, MV_REINSERTION = 4, MV_IDENTITY = 5 };

typedef struct orc_gcmc_cfg {
    double F[9], c2f[9];          /* simulation box, column-major */
    double r2_cut, omega, rep[3];
    int64_t n_atoms;
    const double* atom_xf;        /* n_atoms x 3, fractional */
    const int32_t* atom_type;     /* force-field species */
    int32_t n_types;              /* force-field species count */
    const double* eps;            /* [n_types * n_types] */
    const double* sig2;
    int32_t n_species, capacity;
    int32_t n_sites[GC_NS], rotatable[GC_NS], site_type[GC_NS][GC_MS];
    double off[GC_NS][GC_MS][3];
    double cum[GC_NM];
    int64_t n_burn, n_sample, adjust_every, cycles_per_block;
    int32_t sample_frequency, snapshot_frequency;
    int32_t dens_species, dens_sim_box;
    uint32_t dens_mask;
    int32_t dens_n[3];
    double temperature, fug[GC_NS];
    uint64_t seed;
    uint32_t chain;
} orc_gcmc_cfg;

typedef struct orc_gcmc_state {
    int64_t outer, t, n_trans_prop, n_trans_acc, num_snapshots;
    int32_t n[GC_NS];
    int32_t block, error;
    double delta, e_gh, e_gg;
    int64_t n_prop[GC_NM], n_acc[GC_NM];
    int64_t blk_samples[GC_NB];
    int64_t blk_n[GC_NB][GC_NS], blk_n2[GC_NB][GC_NS];
    double blk_gh[GC_NB], blk_gg[GC_NB], blk_gh2[GC_NB], blk_gg2[GC_NB], blk_un[GC_NB];
} orc_gcmc_state;

static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

static void uniforms(uint64_t seed, uint32_t chain, uint64_t t, double u[10]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t w = 0; w < 5; ++w) {
        uint32_t ctr[4] = {(uint32_t)t, (uint32_t)(t >> 32), chain, GCMC_TAG | w};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        u[2 * w] = u53(o[0], o[1]);
        u[2 * w + 1] = u53(o[2], o[3]);
    }
}

/* Xtals nearest_image! (one sign subtraction when |dx| > 0.5) */
static inline double nearest(double d) {
    if (fabs(d) > 0.5) d -= (d > 0.0) ? 1.0 : -1.0;
    return d;
}

/* lennard_jones + the pair test of vdw_energy; returns 1 on overlap */
static int pair(const orc_gcmc_cfg* c, const double* x, const double* y, int ti, int tj, double* acc) {
    double d0 = nearest(y[0] - x[0]), d1 = nearest(y[1] - x[1]), d2 = nearest(y[2] - x[2]);
    double c0 = c->F[0] * d0 + c->F[3] * d1 + c->F[6] * d2;
    double c1 = c->F[1] * d0 + c->F[4] * d1 + c->F[7] * d2;
    double c2 = c->F[2] * d0 + c->F[5] * d1 + c->F[8] * d2;
    double r2 = c0 * c0 + c1 * c1 + c2 * c2;
    if (r2 < c->r2_cut) {
        if (r2 < R2_OVERLAP) return 1;
        int64_t p = (int64_t)ti * c->n_types + tj;
        double s = c->sig2[p] / r2;
        double ratio = s * s * s;
        *acc += 4.0 * c->eps[p] * ratio * (ratio - 1.0);
    }
    return 0;
}

#define SITE(s, m, k) (sites + ((((int64_t)(s) * c->capacity + (m)) * GC_MS) + (k)) * 3)
#define COM(s, m) (com + ((int64_t)(s) * c->capacity + (m)) * 3)

/* potential_energy (src/gcmc.jl:111-124): guest-host and guest-guest LJ of
 * a molecule with sites x (species ts) against every stored molecule but
 * (ex_s, ex_m); vdw_energy returns +Inf on the first overlap. */
static void energy(const orc_gcmc_cfg* c, const int32_t* n, const double* sites, const double* x, int ts, int ex_s,
                   int ex_m, double* gh, double* gg) {
    double agh = 0.0, agg = 0.0;
    int ogh = 0, ogg = 0;
    for (int k = 0; k < c->n_sites[ts] && !ogh; ++k)
        for (int64_t a = 0; a < c->n_atoms && !ogh; ++a)
            ogh |= pair(c, x + 3 * k, c->atom_xf + 3 * a, c->site_type[ts][k], c->atom_type[a], &agh);
    for (int s2 = 0; s2 < c->n_species && !ogg; ++s2)
        for (int m2 = 0; m2 < n[s2] && !ogg; ++m2) {
            if (s2 == ex_s && m2 == ex_m) continue;
            for (int k = 0; k < c->n_sites[ts] && !ogg; ++k)
                for (int k2 = 0; k2 < c->n_sites[s2] && !ogg; ++k2)
                    ogg |= pair(c, x + 3 * k, SITE(s2, m2, k2), c->site_type[ts][k], c->site_type[s2][k2], &agg);
        }
    *gh = ogh ? INFINITY : agh;
    *gg = ogg ? INFINITY : agg;
}

/* xf = com + c_to_f (R off) (insert_w_random_orientation!, src/mc_helpers.jl:30-49) */
static void new_molecule(const orc_gcmc_cfg* c, int s, const double com[3], const double* R, int rot, double* x) {
    for (int k = 0; k < c->n_sites[s]; ++k) {
        double o[3] = {c->off[s][k][0], c->off[s][k][1], c->off[s][k][2]};
        if (rot) {
            double r[3];
            for (int i = 0; i < 3; ++i) r[i] = R[i] * o[0] + R[i + 3] * o[1] + R[i + 6] * o[2];
            memcpy(o, r, sizeof(o));
        }
        for (int i = 0; i < 3; ++i)
            x[3 * k + i] = com[i] + (c->c2f[i] * o[0] + c->c2f[i + 3] * o[1] + c->c2f[i + 6] * o[2]);
    }
}

/* c_to_f R f_to_c (random_rotation_matrix(box), src/molecule.jl:246) */
static void frac_rotation(const orc_gcmc_cfg* c, const double* R, double* Rf) {
    double T[9];
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r)
            T[r + 3 * col] = c->c2f[r] * R[3 * col] + c->c2f[r + 3] * R[3 * col + 1] + c->c2f[r + 6] * R[3 * col + 2];
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r)
            Rf[r + 3 * col] = T[r] * c->F[3 * col] + T[r + 3] * c->F[3 * col + 1] + T[r + 6] * c->F[3 * col + 2];
}

static void remove_molecule(const orc_gcmc_cfg* c, int32_t* n, double* sites, double* com, int s, int m) {
    for (int q = m + 1; q < n[s]; ++q) {
        memcpy(SITE(s, q - 1, 0), SITE(s, q, 0), sizeof(double) * 3 * GC_MS);
        memcpy(COM(s, q - 1), COM(s, q), sizeof(double) * 3);
    }
    n[s] -= 1;
}

static void store(const orc_gcmc_cfg* c, double* sites, double* com, int s, int m, const double* x, const double* cm) {
    memcpy(SITE(s, m, 0), x, sizeof(double) * 3 * c->n_sites[s]);
    memcpy(COM(s, m), cm, sizeof(double) * 3);
}

static int64_t voxel(double x, int n) {
    int64_t v = (int64_t)floor(x * (double)n);
    if (v < 0) v += n;
    else if (v >= n) v -= n;
    return v;
}

/* update_density! (src/grid.jl:78-106) */
static void density_update(const orc_gcmc_cfg* c, const int32_t* n, const double* sites, const double* com,
                           double* dens) {
    int s = c->dens_species;
    for (int m = 0; m < n[s]; ++m) {
        if (!c->dens_sim_box) {
            const double* cm = COM(s, m);
            double c0 = cm[0] * c->rep[0], c1 = cm[1] * c->rep[1], c2 = cm[2] * c->rep[2];
            if (c0 > 1.0 || c1 > 1.0 || c2 > 1.0 || c0 < 0.0 || c1 < 0.0 || c2 < 0.0) continue;
        }
        for (int k = 0; k < c->n_sites[s]; ++k) {
            if (!((c->dens_mask >> k) & 1u)) continue;
            const double* x = SITE(s, m, k);
            double y[3] = {x[0], x[1], x[2]};
            if (!c->dens_sim_box)
                for (int i = 0; i < 3; ++i) y[i] *= c->rep[i];
            int64_t v = voxel(y[0], c->dens_n[0]) +
                        (int64_t)c->dens_n[0] * (voxel(y[1], c->dens_n[1]) + (int64_t)c->dens_n[1] * voxel(y[2], c->dens_n[2]));
            dens[v] += 1.0;
        }
    }
}

/* System energy from scratch (total_vdw_energy, src/vdw_energetics.jl:146-193). */
void orc_gcmc_total_energy(const orc_gcmc_cfg* c, const int32_t* n, const double* sites, double* gh, double* gg) {
    double tgh = 0.0, tgg = 0.0;
    for (int s = 0; s < c->n_species; ++s)
        for (int m = 0; m < n[s]; ++m) {
            double a, b;
            energy(c, n, sites, SITE(s, m, 0), s, s, m, &a, &b);
            tgh += a;
            tgg += b;
        }
    *gh = tgh;
    *gg = tgg / 2.0;
}

/* Advances the chain by n_cycles outer cycles (the body of src/gcmc.jl:455-680). */
int32_t orc_gcmc_run(const orc_gcmc_cfg* c, int64_t n_cycles, orc_gcmc_state* st, double* sites, double* com,
                     double* dens) {
    const double T = c->temperature;
    const int S = c->n_species;
    int32_t* n = st->n;
    double x_old[GC_MS * 3], x_new[GC_MS * 3];
    for (int64_t cyc = 0; cyc < n_cycles && !st->error; ++cyc) {
        st->outer += 1;
        int64_t n_tot = 0;
        for (int s = 0; s < S; ++s) n_tot += n[s];
        for (int64_t it = 0; it < n_tot + 1; ++it) {
            st->t += 1;
            double u[10];
            uniforms(c->seed, c->chain, (uint64_t)st->t, u);
            int s = S > 1 ? (int)(u[0] * (double)S) : 0;
            if (s > S - 1) s = S - 1;
            int n_i = n[s];
            int mv = 0;
            while (mv < GC_NM - 1 && !(u[1] < c->cum[mv])) ++mv;
            st->n_prop[mv] += 1;
            int rot = c->rotatable[s] != 0;
            int m = 0;
            if (n_i > 0) {
                m = (int)(u[2] * (double)n_i);
                if (m > n_i - 1) m = n_i - 1;
            }
            if (mv == MV_INSERTION) {
                if (n_i == c->capacity) {
                    st->error = 1;
                    break;
                }
                double R[9], cm[3] = {u[3], u[4], u[5]}, gh, gg;
                if (rot) orc_arvo_rotation(&u[6], R);
                new_molecule(c, s, cm, R, rot, x_new);
                energy(c, n, sites, x_new, s, -1, -1, &gh, &gg);
                double dE = (gg + 0.0) + gh + 0.0;
                double acc = c->fug[s] * c->omega / ((double)(n_i + 1) * KB * T) * exp(-dE / T);
                if (u[9] < acc) {
                    store(c, sites, com, s, n_i, x_new, cm);
                    n[s] = n_i + 1;
                    st->n_acc[mv] += 1;
                    st->e_gh += gh;
                    st->e_gg += gg;
                }
            } else if (mv == MV_DELETION && n_i != 0) {
                double gh, gg;
                energy(c, n, sites, SITE(s, m, 0), s, s, m, &gh, &gg);
                double dE = (gg + 0.0) + gh + 0.0;
                double acc = (double)n_i * KB * T / (c->fug[s] * c->omega) * exp(dE / T);
                if (u[9] < acc) {
                    remove_molecule(c, n, sites, com, s, m);
                    st->n_acc[mv] += 1;
                    st->e_gh -= gh;
                    st->e_gg -= gg;
                }
            } else if ((mv == MV_TRANSLATION || (mv == MV_ROTATION && rot) || mv == MV_REINSERTION) && n_i != 0) {
                if (mv == MV_TRANSLATION) st->n_trans_prop += 1;
                const int ns = c->n_sites[s];
                double c0[3], c1[3], gh0, gg0, gh1, gg1;
                memcpy(c0, COM(s, m), sizeof(c0));
                memcpy(c1, c0, sizeof(c1));
                memcpy(x_old, SITE(s, m, 0), sizeof(double) * 3 * ns);
                memcpy(x_new, x_old, sizeof(double) * 3 * ns);
                energy(c, n, sites, x_old, s, s, m, &gh0, &gg0);
                if (mv == MV_TRANSLATION) {
                    /* random_translation!, src/mc_helpers.jl:142-153 */
                    double d[3] = {st->delta * (u[3] - 0.5), st->delta * (u[4] - 0.5), st->delta * (u[5] - 0.5)};
                    double f[3];
                    for (int i = 0; i < 3; ++i) f[i] = c->c2f[i] * d[0] + c->c2f[i + 3] * d[1] + c->c2f[i + 6] * d[2];
                    for (int k = 0; k < ns; ++k)
                        for (int i = 0; i < 3; ++i) x_new[3 * k + i] += f[i];
                    for (int i = 0; i < 3; ++i) c1[i] += f[i];
                    if (c1[0] > 1.0 || c1[1] > 1.0 || c1[2] > 1.0 || c1[0] < 0.0 || c1[1] < 0.0 || c1[2] < 0.0) {
                        double nc[3], e[3];
                        for (int i = 0; i < 3; ++i) {
                            nc[i] = c1[i] >= 1.0 ? c1[i] - 1.0 : (c1[i] < 0.0 ? c1[i] + 1.0 : c1[i]);
                            e[i] = nc[i] - c1[i];
                        }
                        for (int k = 0; k < ns; ++k)
                            for (int i = 0; i < 3; ++i) x_new[3 * k + i] += e[i];
                        for (int i = 0; i < 3; ++i) c1[i] += e[i];
                    }
                } else {
                    if (mv == MV_REINSERTION) {
                        double e[3] = {u[3] - c0[0], u[4] - c0[1], u[5] - c0[2]};
                        for (int k = 0; k < ns; ++k)
                            for (int i = 0; i < 3; ++i) x_new[3 * k + i] += e[i];
                        for (int i = 0; i < 3; ++i) c1[i] += e[i];
                    }
                    if (rot) {
                        double R[9], Rf[9];
                        orc_arvo_rotation(&u[6], R);
                        frac_rotation(c, R, Rf);
                        for (int k = 0; k < ns; ++k) {
                            double a = x_new[3 * k] - c1[0], b = x_new[3 * k + 1] - c1[1], g = x_new[3 * k + 2] - c1[2];
                            for (int i = 0; i < 3; ++i)
                                x_new[3 * k + i] = Rf[i] * a + Rf[i + 3] * b + Rf[i + 6] * g + c1[i];
                        }
                    }
                }
                energy(c, n, sites, x_new, s, s, m, &gh1, &gg1);
                double e_new = (gg1 + 0.0) + gh1 + 0.0, e_old = (gg0 + 0.0) + gh0 + 0.0;
                if (u[9] < exp(-(e_new - e_old) / T)) {
                    store(c, sites, com, s, m, x_new, c1);
                    st->n_acc[mv] += 1;
                    if (mv == MV_TRANSLATION) st->n_trans_acc += 1;
                    st->e_gh += gh1 - gh0;
                    st->e_gg += gg1 - gg0;
                }
            } else if (mv == MV_IDENTITY && n_i != 0) {
                /* src/gcmc.jl:566-613 */
                const int ns = c->n_sites[s];
                double c0[3], gh0, gg0, gh1, gg1;
                memcpy(c0, COM(s, m), sizeof(c0));
                memcpy(x_old, SITE(s, m, 0), sizeof(double) * 3 * ns);
                energy(c, n, sites, x_old, s, s, m, &gh0, &gg0);
                remove_molecule(c, n, sites, com, s, m);
                int pick = (int)(u[3] * (double)(S - 1));
                if (pick > S - 2) pick = S - 2;
                int s2 = pick < s ? pick : pick + 1;
                int n_j = n[s2];
                if (n_j == c->capacity) {
                    st->error = 1;
                    break;
                }
                double R[9];
                int rot2 = c->rotatable[s2] != 0;
                if (rot2) orc_arvo_rotation(&u[6], R);
                new_molecule(c, s2, c0, R, rot2, x_new);
                energy(c, n, sites, x_new, s2, -1, -1, &gh1, &gg1);
                double e_new = (gg1 + 0.0) + gh1 + 0.0, e_old = (gg0 + 0.0) + gh0 + 0.0;
                double acc = ((double)n_i * c->fug[s2] / ((double)(n_j + 1) * c->fug[s])) * exp(-(e_new - e_old) / T);
                if (u[9] < acc) {
                    store(c, sites, com, s2, n_j, x_new, c0);
                    n[s2] = n_j + 1;
                    st->n_acc[mv] += 1;
                    st->e_gh += gh1 - gh0;
                    st->e_gg += gg1 - gg0;
                } else {
                    store(c, sites, com, s, n[s], x_old, c0);
                    n[s] += 1;
                }
            }
            if (st->outer > c->n_burn && st->t % c->sample_frequency == 0) {
                int b = st->block;
                int64_t ntot = 0;
                for (int q = 0; q < GC_NS; ++q) {
                    ntot += n[q];
                    st->blk_n[b][q] += n[q];
                    st->blk_n2[b][q] += (int64_t)n[q] * n[q];
                }
                st->blk_samples[b] += 1;
                st->blk_gh[b] += st->e_gh;
                st->blk_gg[b] += st->e_gg;
                st->blk_gh2[b] += st->e_gh * st->e_gh;
                st->blk_gg2[b] += st->e_gg * st->e_gg;
                st->blk_un[b] += ((st->e_gg + 0.0) + st->e_gh + 0.0) * (double)ntot;
            }
        }
        if (st->error) break;
        if (st->outer <= c->n_burn && c->adjust_every > 0 && st->outer % c->adjust_every == 0) {
            double rate = (double)st->n_trans_acc / (double)st->n_trans_prop;
            if (rate > 0.5) st->delta = 1.2 * st->delta;
            else if (rate < 0.4) st->delta = st->delta / 1.2;
            st->n_trans_prop = 0;
            st->n_trans_acc = 0;
        }
        if (st->outer > c->n_burn && st->block != GC_NB - 1 && (st->outer - c->n_burn) % c->cycles_per_block == 0)
            st->block += 1;
        if (st->outer > c->n_burn && st->outer % c->snapshot_frequency == 0) {
            if (c->dens_species >= 0 && dens) density_update(c, n, sites, com, dens);
            st->num_snapshots += 1;
        }
    }
    return st->error;
}
