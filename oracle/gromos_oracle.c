/*
 * CPU oracle: plain-C restatement of the gromos-rs nonbonded hot path.
 * TEST INFRASTRUCTURE ONLY (see gromos_oracle.h).  Compile with
 * -ffp-contract=off and without fast-math: operation order below follows the
 * Rust source so that every per-pair value is the reference's value.
 */
#include "gromos_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* boundary conditions: crates/gromos-core/src/math.rs:20-74           */
/* ------------------------------------------------------------------ */
void or_pbc_init(or_pbc *p, int kind, const double *box) {
    p->kind = kind;
    for (int a = 0; a < 3; ++a) {
        p->a[a] = p->b[a] = p->c[a] = 0.0;
        if (kind == 2) {
            p->a[a] = box[a];
            p->b[a] = box[3 + a];
            p->c[a] = box[6 + a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (kind == 2) p->L[a] = a == 0 ? p->a[0] : (a == 1 ? p->b[1] : p->c[2]);
        else p->L[a] = box ? box[a] : 0.0;
        p->invL[a] = (kind == 1) ? 1.0 / p->L[a] : 0.0;   /* DVec3::recip */
        p->half[a] = p->L[a] * 0.5;
    }
}

/* Triclinic::nearest_image (math.rs:105-141): z, then y, then x reduction
 * loops over the lower-triangular box vectors, in the reference's operation
 * order. */
static void nim_triclinic(const or_pbc *p, double *n) {
    const double *a = p->a, *b = p->b, *c = p->c;
    while (n[2] >= 0.5 * c[2]) { n[0] -= c[0]; n[1] -= c[1]; n[2] -= c[2]; }
    while (n[2] <= -0.5 * c[2]) { n[0] += c[0]; n[1] += c[1]; n[2] += c[2]; }
    double ny = n[1] - (c[1] / c[2]) * n[2];
    while (ny >= 0.5 * b[1]) { n[0] -= b[0]; n[1] -= b[1]; n[2] -= b[2]; ny = n[1] - (c[1] / c[2]) * n[2]; }
    while (ny <= -0.5 * b[1]) { n[0] += b[0]; n[1] += b[1]; n[2] += b[2]; ny = n[1] - (c[1] / c[2]) * n[2]; }
    double nx = n[0] - b[0] * (n[1] - c[1] * n[2] / c[2]) / b[1] - c[0] * n[2] / c[2];
    while (nx >= 0.5 * a[0]) {
        n[0] -= a[0]; n[1] -= a[1]; n[2] -= a[2];
        nx = n[0] - b[0] * (n[1] - c[1] * n[2] / c[2]) / b[1] - c[0] * n[2] / c[2];
    }
    while (nx <= -0.5 * a[0]) {
        n[0] += a[0]; n[1] += a[1]; n[2] += a[2];
        nx = n[0] - b[0] * (n[1] - c[1] * n[2] / c[2]) / b[1] - c[0] * n[2] / c[2];
    }
}

/* Rectangular::nearest_image (math.rs:58-68): r = ri - rj; r -= L * round(r * invL).
 * C round() is half-away-from-zero like f64::round. Vacuum (math.rs:24-34): ri - rj. */
void or_nearest_image(const or_pbc *p, const double *ri, const double *rj, double *r) {
    for (int a = 0; a < 3; ++a) {
        double d = ri[a] - rj[a];
        if (p->kind == 1) d = d - p->L[a] * round(d * p->invL[a]);
        r[a] = d;
    }
    if (p->kind == 2) nim_triclinic(p, r);
}

static inline double len2(const double r[3]) { return r[0] * r[0] + r[1] * r[1] + r[2] * r[2]; }

/* ------------------------------------------------------------------ */
/* CRFParameters::new (nonbonded/params.rs:78-97)                      */
/* ------------------------------------------------------------------ */
void or_crf_new(double cutoff, double epsilon, double rf_epsilon, double rf_kappa, or_crf *o) {
    double kappa_cut = rf_kappa * cutoff;
    double kappa_cut2 = kappa_cut * kappa_cut;
    double crf = (2.0 * (epsilon - rf_epsilon) * (1.0 + kappa_cut) - rf_epsilon * kappa_cut2) /
                 ((epsilon + 2.0 * rf_epsilon) * (1.0 + kappa_cut) + rf_epsilon * kappa_cut2);
    double cut3i = 1.0 / (cutoff * cutoff * cutoff);
    double crf_cut3i = crf * cut3i;
    o->crf_cut3i = crf_cut3i;
    o->crf_2cut3i = crf_cut3i / 2.0;
    o->crf_cut = (1.0 - crf / 2.0) / cutoff;
    o->cutoff_sq = cutoff * cutoff;
}

/* ------------------------------------------------------------------ */
/* lj_crf_interaction (nonbonded/innerloops.rs:22-55)                  */
/* ------------------------------------------------------------------ */
void or_lj_crf_interaction(const double r[3], double c6, double c12, double q_prod,
                           const or_crf *crf, double *f, double *e_lj, double *e_crf) {
    double r2 = len2(r);
    if (r2 < 1e-10) { *f = 0.0; *e_lj = 0.0; *e_crf = 0.0; return; }
    double inv_r2 = 1.0 / r2;
    double inv_r6 = inv_r2 * inv_r2 * inv_r2;
    double elj = (c12 * inv_r6 - c6) * inv_r6;
    double flj = (12.0 * c12 * inv_r6 - 6.0 * c6) * inv_r6 * inv_r2;
    double inv_r = sqrt(inv_r2);
    double ecrf = q_prod * (inv_r - crf->crf_2cut3i * r2 - crf->crf_cut);
    double fcrf = q_prod * (inv_r * inv_r2 + crf->crf_cut3i);
    *f = flj + fcrf;
    *e_lj = elj;
    *e_crf = ecrf;
}

/* ------------------------------------------------------------------ */
/* pair lists                                                          */
/* ------------------------------------------------------------------ */
static void pl_push(or_plist *l, uint32_t i, uint32_t j) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 1024;
        l->p = (uint32_t *)realloc(l->p, (size_t)l->cap * 2 * sizeof(uint32_t));
    }
    l->p[2 * l->n] = i;
    l->p[2 * l->n + 1] = j;
    l->n++;
}

void or_pairlist_free(or_pairlist *pl) {
    free(pl->solute_short.p); free(pl->solute_long.p);
    free(pl->solvent_short.p); free(pl->solvent_long.p);
    memset(pl, 0, sizeof(*pl));
}

void or_pairlist_clear(or_pairlist *pl) {
    pl->solute_short.n = pl->solute_long.n = pl->solvent_short.n = pl->solvent_long.n = 0;
}

static int bsearch_i32(const int32_t *a, int32_t n, int32_t key) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == key;
}

static int contains_i32(const int32_t *a, int32_t n, int32_t key) {
    for (int32_t k = 0; k < n; ++k) if (a[k] == key) return 1;
    return 0;
}

/* topology.rs:853-865: binary search in both exclusion lists, linear search in 1-4 lists */
int or_is_excluded_or_14(const or_topo *t, int i, int j) {
    if (bsearch_i32(t->excl + t->excl_off[i], t->excl_off[i + 1] - t->excl_off[i], j)) return 1;
    if (bsearch_i32(t->excl + t->excl_off[j], t->excl_off[j + 1] - t->excl_off[j], i)) return 1;
    if (t->o14_off) {
        if (contains_i32(t->o14 + t->o14_off[i], t->o14_off[i + 1] - t->o14_off[i], j)) return 1;
        if (contains_i32(t->o14 + t->o14_off[j], t->o14_off[j + 1] - t->o14_off[j], i)) return 1;
    }
    return 0;
}

/* ChargeGroup::center_of_geometry (topology.rs:408-416): sum from zero, then divide */
static void cog(const or_topo *t, const double *pos, int cg, double *c) {
    int s = t->cg_start[cg], e = t->cg_start[cg + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int a = s; a < e; ++a) { sx += pos[3 * a]; sy += pos[3 * a + 1]; sz += pos[3 * a + 2]; }
    double n = (double)(e - s);
    c[0] = sx / n; c[1] = sy / n; c[2] = sz / n;
}

/* classification of one CG pair (cg1 < cg2), shared by the standard and the
 * cell-list variants: pairlist.rs:252-301 (solute-solute), :304-350
 * (solute-solvent), :359-415 (solvent-solvent); pairlist.rs:698-777 restates
 * the same cases for the cell list. */
static void process_cg_pair(const or_topo *t, const double *pos, const or_pbc *pbc,
                            const double *cogs, int cg1, int cg2, double c2s, double c2l,
                            or_pairlist *out) {
    double r[3];
    int nsc = t->n_solute_cg;
    const int32_t *cs = t->cg_start;
    if (cg2 < nsc) {
        or_nearest_image(pbc, cogs + 3 * cg1, cogs + 3 * cg2, r);
        double d2 = len2(r);
        if (d2 > c2l) return;
        if (d2 > c2s) {
            for (int a1 = cs[cg1]; a1 < cs[cg1 + 1]; ++a1)
                for (int a2 = cs[cg2]; a2 < cs[cg2 + 1]; ++a2) pl_push(&out->solute_long, a1, a2);
            return;
        }
        for (int a1 = cs[cg1]; a1 < cs[cg1 + 1]; ++a1)
            for (int a2 = cs[cg2]; a2 < cs[cg2 + 1]; ++a2)
                if (!or_is_excluded_or_14(t, a1, a2)) pl_push(&out->solute_short, a1, a2);
    } else if (cg1 < nsc) {
        int fa = cs[cg2];
        or_nearest_image(pbc, cogs + 3 * cg1, pos + 3 * fa, r);
        double d2 = len2(r);
        if (d2 > c2l) return;
        or_plist *dst = (d2 > c2s) ? &out->solute_long : &out->solute_short;
        for (int a1 = cs[cg1]; a1 < cs[cg1 + 1]; ++a1)
            for (int a2 = cs[cg2]; a2 < cs[cg2 + 1]; ++a2) pl_push(dst, a1, a2);
    } else {
        int f1 = cs[cg1], f2 = cs[cg2];
        or_nearest_image(pbc, pos + 3 * f1, pos + 3 * f2, r);
        double d2 = len2(r);
        if (d2 > c2l) return;
        if (d2 > c2s) {
            for (int a1 = cs[cg1]; a1 < cs[cg1 + 1]; ++a1)
                for (int a2 = cs[cg2]; a2 < cs[cg2 + 1]; ++a2) pl_push(&out->solvent_long, a1, a2);
            return;
        }
        pl_push(&out->solvent_short, f1, f2);
    }
}

static double *solute_cogs(const or_topo *t, const double *pos) {
    double *c = (double *)malloc(sizeof(double) * 3 * (size_t)(t->n_solute_cg > 0 ? t->n_solute_cg : 1));
    for (int cg = 0; cg < t->n_solute_cg; ++cg) cog(t, pos, cg, c + 3 * cg);
    return c;
}

static void intra_cg_pairs(const or_topo *t, int cg, or_pairlist *out) {
    /* pairlist.rs:238-249: non-excluded intra-CG pairs, always short */
    for (int a1 = t->cg_start[cg]; a1 < t->cg_start[cg + 1]; ++a1)
        for (int a2 = a1 + 1; a2 < t->cg_start[cg + 1]; ++a2)
            if (!or_is_excluded_or_14(t, a1, a2)) pl_push(&out->solute_short, a1, a2);
}

/* StandardPairlistAlgorithm::update_chargegroup_based (pairlist.rs:183-442).
 * cutoff2 = (rc + skin).powi(2) (pairlist.rs:192-193); loop order as the source. */
void or_pairlist_standard(const or_topo *t, const double *pos, const or_pbc *pbc, double rcutp,
                          double rcutl, double skin, or_pairlist *out) {
    or_pairlist_clear(out);
    double s = rcutp + skin, l = rcutl + skin;
    double c2s = s * s, c2l = l * l;
    double *cogs = solute_cogs(t, pos);
    int nsc = t->n_solute_cg, ncg = t->n_cg;
    for (int cg1 = 0; cg1 < nsc; ++cg1) {
        intra_cg_pairs(t, cg1, out);
        for (int cg2 = cg1 + 1; cg2 < nsc; ++cg2) process_cg_pair(t, pos, pbc, cogs, cg1, cg2, c2s, c2l, out);
        for (int cg2 = nsc; cg2 < ncg; ++cg2) process_cg_pair(t, pos, pbc, cogs, cg1, cg2, c2s, c2l, out);
    }
    for (int cg1 = nsc; cg1 < ncg; ++cg1)
        for (int cg2 = cg1 + 1; cg2 < ncg; ++cg2) process_cg_pair(t, pos, pbc, cogs, cg1, cg2, c2s, c2l, out);
    free(cogs);
}

/* CellListPairlistAlgorithm (pairlist.rs:537-655): grid of floor(L/(rcl+skin)) >= 1
 * cells per axis (:657-663), cell of the reference point by floor(pos/cell) wrapped
 * (:666-671), de-duplicated 27-neighbourhood (:675-692). Falls back to the standard
 * algorithm for non-rectangular boxes (:521-532). */
static long rem_euclid(long a, long m) { long r = a % m; return r < 0 ? r + m : r; }

void or_pairlist_cell(const or_topo *t, const double *pos, const or_pbc *pbc, double rcutp,
                      double rcutl, double skin, or_pairlist *out) {
    if (pbc->kind != 1) { or_pairlist_standard(t, pos, pbc, rcutp, rcutl, skin, out); return; }
    or_pairlist_clear(out);
    int ncg = t->n_cg, nsc = t->n_solute_cg;
    if (ncg == 0) return;
    double s = rcutp + skin, l = rcutl + skin;
    double c2s = s * s, c2l = l * l;
    double minc = rcutl + skin;
    double *cogs = solute_cogs(t, pos);
    for (int cg = 0; cg < nsc; ++cg) intra_cg_pairs(t, cg, out);
    long gd[3];
    double csz[3];
    for (int a = 0; a < 3; ++a) {
        long g = 1;
        if (minc > 0.0 && pbc->L[a] > 0.0) { g = (long)floor(pbc->L[a] / minc); if (g < 1) g = 1; }
        gd[a] = g;
        csz[a] = pbc->L[a] / (double)g;
    }
    long ncell = gd[0] * gd[1] * gd[2];
    int *cell_of = (int *)malloc(sizeof(int) * ncg);
    int *count = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (int cg = 0; cg < ncg; ++cg) {
        const double *rp = (cg < nsc) ? cogs + 3 * cg : pos + 3 * t->cg_start[cg];
        long cx = rem_euclid((long)floor(rp[0] / csz[0]), gd[0]);
        long cy = rem_euclid((long)floor(rp[1] / csz[1]), gd[1]);
        long cz = rem_euclid((long)floor(rp[2] / csz[2]), gd[2]);
        cell_of[cg] = (int)(cx + cy * gd[0] + cz * gd[0] * gd[1]);
        count[cell_of[cg] + 1]++;
    }
    for (long c = 0; c < ncell; ++c) count[c + 1] += count[c];
    int *members = (int *)malloc(sizeof(int) * ncg);
    int *fill = (int *)malloc(sizeof(int) * ncell);
    for (long c = 0; c < ncell; ++c) fill[c] = count[c];
    for (int cg = 0; cg < ncg; ++cg) members[fill[cell_of[cg]]++] = cg;
    for (long ca = 0; ca < ncell; ++ca) {
        if (count[ca + 1] == count[ca]) continue;
        long cx = ca % gd[0], cy = (ca / gd[0]) % gd[1], cz = ca / (gd[0] * gd[1]);
        long nb[27];
        int nn = 0;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    long c = rem_euclid(cx + dx, gd[0]) + rem_euclid(cy + dy, gd[1]) * gd[0] +
                             rem_euclid(cz + dz, gd[2]) * gd[0] * gd[1];
                    int dup = 0;
                    for (int k = 0; k < nn; ++k) if (nb[k] == c) { dup = 1; break; }
                    if (!dup) nb[nn++] = c;
                }
        for (int k = 0; k < nn; ++k) {
            long cb = nb[k];
            if (cb < ca || count[cb + 1] == count[cb]) continue;
            if (cb == ca) {
                for (int x = count[ca]; x < count[ca + 1]; ++x)
                    for (int y = x + 1; y < count[ca + 1]; ++y) {
                        int a = members[x], b = members[y];
                        process_cg_pair(t, pos, pbc, cogs, a < b ? a : b, a < b ? b : a, c2s, c2l, out);
                    }
            } else {
                for (int x = count[ca]; x < count[ca + 1]; ++x)
                    for (int y = count[cb]; y < count[cb + 1]; ++y) {
                        int a = members[x], b = members[y];
                        process_cg_pair(t, pos, pbc, cogs, a < b ? a : b, a < b ? b : a, c2s, c2l, out);
                    }
            }
        }
    }
    free(cell_of); free(count); free(members); free(fill); free(cogs);
}

/* ------------------------------------------------------------------ */
/* ForceStorage (params.rs:99-138)                                     */
/* ------------------------------------------------------------------ */
void or_storage_init(or_storage *s, int n) {
    s->n = n;
    s->f = (double *)calloc((size_t)3 * (n > 0 ? n : 1), sizeof(double));
    s->e_lj = s->e_crf = 0.0;
    memset(s->vir, 0, sizeof(s->vir));
}
void or_storage_free(or_storage *s) { free(s->f); s->f = NULL; }
void or_storage_clear(or_storage *s) {
    memset(s->f, 0, sizeof(double) * 3 * (size_t)s->n);
    s->e_lj = s->e_crf = 0.0;
    memset(s->vir, 0, sizeof(s->vir));
}
void or_storage_merge(or_storage *d, const or_storage *s) {
    for (int64_t k = 0; k < 3 * (int64_t)d->n; ++k) d->f[k] += s->f[k];
    d->e_lj += s->e_lj;
    d->e_crf += s->e_crf;
    for (int k = 0; k < 9; ++k) d->vir[k] += s->vir[k];
}

static inline void accum_pair(or_storage *s, int i, int j, const double r[3], double fm,
                              double elj, double ecrf, int virial) {
    double fx = r[0] * fm, fy = r[1] * fm, fz = r[2] * fm;
    s->f[3 * i] += fx; s->f[3 * i + 1] += fy; s->f[3 * i + 2] += fz;
    s->f[3 * j] -= fx; s->f[3 * j + 1] -= fy; s->f[3 * j + 2] -= fz;
    s->e_lj += elj;
    s->e_crf += ecrf;
    if (virial) {
        double f[3] = {fx, fy, fz};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s->vir[3 * a + b] += r[a] * f[b];
    }
}

/* process_pairs (innerloops.rs:169-210): per-pair nearest image, q = qi*qj*fpepsi */
void or_lj_crf_innerloop(const double *pos, const double *q, const int32_t *iac,
                         const uint32_t *pairs, int64_t np, const double *lj, int32_t nt,
                         const or_crf *crf, const or_pbc *pbc, double fpepsi, int virial,
                         or_storage *s) {
    for (int64_t k = 0; k < np; ++k) {
        int i = (int)pairs[2 * k], j = (int)pairs[2 * k + 1];
        double r[3];
        or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
        const double *p = lj + 4 * ((size_t)iac[i] * nt + iac[j]);
        double qp = q[i] * q[j] * fpepsi;
        double fm, elj, ecrf;
        or_lj_crf_interaction(r, p[0], p[1], qp, crf, &fm, &elj, &ecrf);
        accum_pair(s, i, j, r, fm, elj, ecrf, virial);
    }
}

/* process_pairs_cg_grouped (innerloops.rs:105-162): shift t = nim(ri,rj) - ri + rj
 * from the groups' reference atoms, r = (pos_i + t) - pos_j */
void or_lj_crf_innerloop_cg_grouped(const double *pos, const double *q, const int32_t *iac,
                                    const uint32_t *pairs, const uint32_t *groups, int64_t ng,
                                    const double *lj, int32_t nt, const or_crf *crf,
                                    const or_pbc *pbc, double fpepsi, int virial, or_storage *s) {
    for (int64_t g = 0; g < ng; ++g) {
        const uint32_t *G = groups + 4 * g;
        const double *ri = pos + 3 * G[0], *rj = pos + 3 * G[1];
        double rr[3], tt[3];
        or_nearest_image(pbc, ri, rj, rr);
        for (int a = 0; a < 3; ++a) tt[a] = rr[a] - ri[a] + rj[a];
        for (uint32_t k = G[2]; k < G[3]; ++k) {
            int i = (int)pairs[2 * k], j = (int)pairs[2 * k + 1];
            double r[3];
            for (int a = 0; a < 3; ++a) r[a] = pos[3 * i + a] + tt[a] - pos[3 * j + a];
            const double *p = lj + 4 * ((size_t)iac[i] * nt + iac[j]);
            double qp = q[i] * q[j] * fpepsi;
            double fm, elj, ecrf;
            or_lj_crf_interaction(r, p[0], p[1], qp, crf, &fm, &elj, &ecrf);
            accum_pair(s, i, j, r, fm, elj, ecrf, virial);
        }
    }
}

/* process_solvent_pairs (innerloops.rs:218-309): 4x4 tables from the first pair's
 * molecule, shift from the atom-0 pair, n x n atom pairs per sentinel */
void or_solvent_innerloop(const double *pos, const double *q, const int32_t *iac,
                          const uint32_t *pairs, int64_t np, const double *lj, int32_t nt,
                          const or_crf *crf, const or_pbc *pbc, int32_t n, double fpepsi,
                          int virial, or_storage *s) {
    if (np == 0) return;
    if (n > 4) abort(); /* innerloops.rs:237-240 */
    int first = (int)pairs[0];
    double c6[4][4], c12[4][4], qq[4][4];
    for (int ai = 0; ai < n; ++ai) {
        double qi = q[first + ai];
        for (int aj = 0; aj < n; ++aj) {
            const double *p = lj + 4 * ((size_t)iac[first + ai] * nt + iac[first + aj]);
            c6[ai][aj] = p[0];
            c12[ai][aj] = p[1];
            qq[ai][aj] = qi * q[first + aj] * fpepsi;
        }
    }
    for (int64_t k = 0; k < np; ++k) {
        int i0 = (int)pairs[2 * k], j0 = (int)pairs[2 * k + 1];
        const double *pi0 = pos + 3 * i0, *pj0 = pos + 3 * j0;
        double rf[3], tt[3];
        or_nearest_image(pbc, pi0, pj0, rf);
        for (int a = 0; a < 3; ++a) tt[a] = rf[a] - pi0[a] + pj0[a];
        for (int ai = 0; ai < n; ++ai) {
            int i = i0 + ai;
            double xi[3];
            for (int a = 0; a < 3; ++a) xi[a] = pos[3 * i + a] + tt[a];
            for (int aj = 0; aj < n; ++aj) {
                int j = j0 + aj;
                double r[3];
                for (int a = 0; a < 3; ++a) r[a] = xi[a] - pos[3 * j + a];
                double fm, elj, ecrf;
                or_lj_crf_interaction(r, c6[ai][aj], c12[ai][aj], qq[ai][aj], crf, &fm, &elj, &ecrf);
                accum_pair(s, i, j, r, fm, elj, ecrf, virial);
            }
        }
    }
}

/* rayon fold/reduce restatements (innerloops.rs:312, 612-667, 742-864):
 * below 2048 pairs serial; otherwise chunks of 1024 (plain) / 512 (solvent)
 * folded into per-thread N-sized storages and merged. */
#define OR_PARALLEL_THRESHOLD 2048

void or_lj_crf_innerloop_parallel(const double *pos, const double *q, const int32_t *iac,
                                  const uint32_t *pairs, int64_t np, const double *lj,
                                  int32_t nt, const or_crf *crf, const or_pbc *pbc,
                                  double fpepsi, int virial, or_storage *out) {
    if (np < OR_PARALLEL_THRESHOLD) {
        or_lj_crf_innerloop(pos, q, iac, pairs, np, lj, nt, crf, pbc, fpepsi, virial, out);
        return;
    }
    int nth = or_num_threads();
    or_storage *loc = (or_storage *)malloc(sizeof(or_storage) * nth);
    for (int t = 0; t < nth; ++t) or_storage_init(&loc[t], out->n);
    int64_t nchunk = (np + 1023) / 1024;
#pragma omp parallel num_threads(nth)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
#pragma omp for schedule(dynamic, 4)
        for (int64_t c = 0; c < nchunk; ++c) {
            int64_t b = c * 1024, e = b + 1024 < np ? b + 1024 : np;
            or_lj_crf_innerloop(pos, q, iac, pairs + 2 * b, e - b, lj, nt, crf, pbc, fpepsi, virial, &loc[tid]);
        }
    }
    for (int t = 0; t < nth; ++t) { or_storage_merge(out, &loc[t]); or_storage_free(&loc[t]); }
    free(loc);
}

void or_solvent_innerloop_parallel(const double *pos, const double *q, const int32_t *iac,
                                   const uint32_t *pairs, int64_t np, const double *lj,
                                   int32_t nt, const or_crf *crf, const or_pbc *pbc,
                                   int32_t n, double fpepsi, int virial, or_storage *out) {
    if (np < OR_PARALLEL_THRESHOLD) {
        or_solvent_innerloop(pos, q, iac, pairs, np, lj, nt, crf, pbc, n, fpepsi, virial, out);
        return;
    }
    int nth = or_num_threads();
    or_storage *loc = (or_storage *)malloc(sizeof(or_storage) * nth);
    for (int t = 0; t < nth; ++t) or_storage_init(&loc[t], out->n);
    int64_t nchunk = (np + 511) / 512;
#pragma omp parallel num_threads(nth)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
#pragma omp for schedule(dynamic, 4)
        for (int64_t c = 0; c < nchunk; ++c) {
            int64_t b = c * 512, e = b + 512 < np ? b + 512 : np;
            /* tables come from the first pair of the whole list; all solvent
             * molecules are copies of one template so any chunk gives the same */
            or_solvent_innerloop(pos, q, iac, pairs + 2 * b, e - b, lj, nt, crf, pbc, n, fpepsi, virial, &loc[tid]);
        }
    }
    for (int t = 0; t < nth; ++t) { or_storage_merge(out, &loc[t]); or_storage_free(&loc[t]); }
    free(loc);
}

static void or_cg_grouped_parallel(const double *pos, const double *q, const int32_t *iac,
                                   const uint32_t *pairs, const uint32_t *groups, int64_t ng,
                                   const double *lj, int32_t nt, const or_crf *crf,
                                   const or_pbc *pbc, double fpepsi, int virial, or_storage *out) {
    /* innerloops.rs:490-549: serial below 64 groups, else chunks of groups/threads */
    if (ng < 64) {
        or_lj_crf_innerloop_cg_grouped(pos, q, iac, pairs, groups, ng, lj, nt, crf, pbc, fpepsi, virial, out);
        return;
    }
    int nth = or_num_threads();
    int64_t chunk = ng / nth;
    if (chunk < 1) chunk = 1;
    int64_t nchunk = (ng + chunk - 1) / chunk;
    or_storage *loc = (or_storage *)malloc(sizeof(or_storage) * nth);
    for (int t = 0; t < nth; ++t) or_storage_init(&loc[t], out->n);
#pragma omp parallel num_threads(nth)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < nchunk; ++c) {
            int64_t b = c * chunk, e = b + chunk < ng ? b + chunk : ng;
            or_lj_crf_innerloop_cg_grouped(pos, q, iac, pairs, groups + 4 * b, e - b, lj, nt, crf, pbc, fpepsi, virial, &loc[tid]);
        }
    }
    for (int t = 0; t < nth; ++t) { or_storage_merge(out, &loc[t]); or_storage_free(&loc[t]); }
    free(loc);
}

/* ------------------------------------------------------------------ */
/* rf_excluded_interactions (nonbonded/excluded.rs:25-112)             */
/* ------------------------------------------------------------------ */
void or_rf_excluded_interactions(const or_topo *t, const double *pos, const or_crf *crf,
                                 const or_pbc *pbc, double fpepsi, or_storage *s) {
    const double *q = t->charge;
    for (int i = 0; i < t->n_solute_atoms; ++i) {
        double term = -0.5 * q[i] * q[i] * fpepsi * crf->crf_cut;
        s->e_crf += term;
        for (int k = t->excl_off[i]; k < t->excl_off[i + 1]; ++k) {
            int j = t->excl[k];
            if (j <= i) continue;
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
            double r2 = len2(r);
            double qp = q[i] * q[j] * fpepsi;
            s->e_crf += qp * (-crf->crf_2cut3i * r2 - crf->crf_cut);
            double fs = qp * crf->crf_cut3i;
            double f[3] = {r[0] * fs, r[1] * fs, r[2] * fs};
            for (int a = 0; a < 3; ++a) { s->f[3 * i + a] += f[a]; s->f[3 * j + a] -= f[a]; }
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) s->vir[3 * a + b] += r[a] * f[b];
        }
    }
    for (int i = t->n_solute_atoms; i < t->n_atoms; ++i) {
        for (int k = t->excl_off[i]; k < t->excl_off[i + 1]; ++k) {
            int j = t->excl[k];
            if (j <= i) continue;
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
            double r2 = len2(r);
            s->e_crf += -q[i] * q[j] * fpepsi * crf->crf_2cut3i * r2;
        }
    }
}

/* one_four_interaction_loop (excluded.rs:131-206): cs6/cs12 LJ + CRF, j > i, no cutoff */
void or_one_four_interaction_loop(const or_topo *t, const double *pos, const or_crf *crf,
                                  const or_pbc *pbc, double fpepsi, double cs, or_storage *s) {
    if (!t->o14_off) return;
    double elj14 = 0.0, ecrf14 = 0.0;
    for (int i = 0; i < t->n_atoms; ++i) {
        for (int k = t->o14_off[i]; k < t->o14_off[i + 1]; ++k) {
            int j = t->o14[k];
            if (j <= i) continue;
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
            double r2 = len2(r);
            if (r2 < 1e-10) continue;
            const double *p = t->lj + 4 * ((size_t)t->iac[i] * t->n_types + t->iac[j]);
            double inv_r2 = 1.0 / r2;
            double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            double elj = (p[3] * inv_r6 - p[2]) * inv_r6;
            double flj = (12.0 * p[3] * inv_r6 - 6.0 * p[2]) * inv_r6 * inv_r2;
            double qp = t->charge[i] * t->charge[j] * fpepsi;
            double inv_r = sqrt(inv_r2);
            double ecrf = qp * (inv_r * cs - crf->crf_2cut3i * r2 - crf->crf_cut);
            double fcrf = qp * (inv_r * cs * inv_r2 + crf->crf_cut3i);
            double ft = flj + fcrf;
            double f[3] = {r[0] * ft, r[1] * ft, r[2] * ft};
            for (int a = 0; a < 3; ++a) { s->f[3 * i + a] += f[a]; s->f[3 * j + a] -= f[a]; }
            elj14 += elj;
            ecrf14 += ecrf;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) s->vir[3 * a + b] += r[a] * f[b];
        }
    }
    s->e_lj += elj14;
    s->e_crf += ecrf14;
}

/* ------------------------------------------------------------------ */
/* perturbation (nonbonded/perturbed.rs)                               */
/* ------------------------------------------------------------------ */
/* f64::powi with a run-time exponent: square-and-multiply as LLVM's powi */
static double powi_(double x, int n) {
    int neg = n < 0;
    unsigned u = (unsigned)(neg ? -n : n);
    double r = 1.0;
    for (;;) {
        if (u & 1u) r *= x;
        u >>= 1;
        if (!u) break;
        x *= x;
    }
    return neg ? 1.0 / r : r;
}

/* PerturbedLambdaParams::from_lambda (perturbed.rs:57-101) */
void or_lambda_from(double lj, double ljs, double crfl, double crfs, double ljd, double ljsd,
                    double crfd, double crfsd, int n, or_lambda *o) {
    o->a_lj_lambda_n = powi_(1.0 - lj, n);
    o->a_crf_lambda_n = powi_(1.0 - crfl, n);
    o->b_lj_lambda_n = powi_(lj, n);
    o->b_crf_lambda_n = powi_(crfl, n);
    o->a_lj_lambda_n_1 = powi_(1.0 - lj, n - 1) * ljd;
    o->a_crf_lambda_n_1 = powi_(1.0 - crfl, n - 1) * crfd;
    o->b_lj_lambda_n_1 = powi_(lj, n - 1) * ljd;
    o->b_crf_lambda_n_1 = powi_(crfl, n - 1) * crfd;
    o->a_ljs_lambda = (1.0 - ljs) * ljsd;
    o->a_crfs_lambda = (1.0 - crfs) * crfsd;
    o->b_ljs_lambda = ljs * ljsd;
    o->b_crfs_lambda = crfs * crfsd;
    o->a_ljs_lambda2 = (1.0 - ljs) * (1.0 - ljs);
    o->a_crfs_lambda2 = (1.0 - crfs) * (1.0 - crfs);
    o->b_ljs_lambda2 = ljs * ljs;
    o->b_crfs_lambda2 = crfs * crfs;
    o->lambda_exp = n;
}

/* perturbed_lj_crf_interaction (perturbed.rs:124-266); the soft cutoff term divides
 * crf_2cut3i by sqrt(cut2soft^3) exactly as written there (:190-191). */
void or_perturbed_lj_crf_interaction(const double r[3], double a_c6, double a_c12, double b_c6,
                                     double b_c12, double a_q, double b_q, double alpha_lj,
                                     double alpha_crf, const or_lambda *lp, const or_crf *crf,
                                     double fpepsi, double out[5]) {
    double r2 = len2(r);
    if (r2 < 1e-10) { for (int k = 0; k < 5; ++k) out[k] = 0.0; return; }
    double r6 = r2 * r2 * r2;
    double r4 = r2 * r2;
    double a_c126 = fabs(a_c6) > 1e-10 ? a_c12 / a_c6 : 0.0;
    double b_c126 = fabs(b_c6) > 1e-10 ? b_c12 / b_c6 : 0.0;

    double a_dist2soft = r2 + alpha_crf * lp->b_crfs_lambda2;
    double a_distisoft = 1.0 / sqrt(a_dist2soft);
    double a_dist3isoft = a_distisoft / a_dist2soft;
    double a_dist6soft = r6 + alpha_lj * lp->b_ljs_lambda2 * a_c126;
    double a_dist6isoft = 1.0 / a_dist6soft;

    double b_dist2soft = r2 + alpha_crf * lp->a_crfs_lambda2;
    double b_distisoft = 1.0 / sqrt(b_dist2soft);
    double b_dist3isoft = b_distisoft / b_dist2soft;
    double b_dist6soft = r6 + alpha_lj * lp->a_ljs_lambda2 * b_c126;
    double b_dist6isoft = 1.0 / b_dist6soft;

    double cut2 = crf->cutoff_sq;
    double a_cut2soft = cut2 + alpha_crf * lp->b_crfs_lambda2;
    double b_cut2soft = cut2 + alpha_crf * lp->a_crfs_lambda2;
    double a_cut2soft3 = a_cut2soft * a_cut2soft * a_cut2soft;
    double b_cut2soft3 = b_cut2soft * b_cut2soft * b_cut2soft;
    double a_crf_2cut3i = crf->crf_2cut3i / sqrt(a_cut2soft3);
    double b_crf_2cut3i = crf->crf_2cut3i / sqrt(b_cut2soft3);
    double a_crf_cut3i = 2.0 * a_crf_2cut3i;
    double b_crf_cut3i = 2.0 * b_crf_2cut3i;
    double a_crf_pert = 3.0 * a_crf_2cut3i / a_cut2soft;
    double b_crf_pert = 3.0 * b_crf_2cut3i / b_cut2soft;

    double force = (lp->a_crf_lambda_n * a_q * (a_dist3isoft + a_crf_cut3i) +
                    lp->b_crf_lambda_n * b_q * (b_dist3isoft + b_crf_cut3i)) * fpepsi;
    force += -6.0 * (lp->a_lj_lambda_n * a_c6 * a_dist6isoft * a_dist6isoft +
                     lp->b_lj_lambda_n * b_c6 * b_dist6isoft * b_dist6isoft) * r4;
    force += 12.0 * (lp->a_lj_lambda_n * a_c12 * a_dist6isoft * a_dist6isoft * a_dist6isoft +
                     lp->b_lj_lambda_n * b_c12 * b_dist6isoft * b_dist6isoft * b_dist6isoft) * r4;

    double a_e_lj = (a_c12 * a_dist6isoft - a_c6) * a_dist6isoft;
    double b_e_lj = (b_c12 * b_dist6isoft - b_c6) * b_dist6isoft;
    double crf_cut = crf->crf_cut;
    double a_e_crf = a_q * (a_distisoft - a_crf_2cut3i * r2 - crf_cut);
    double b_e_crf = b_q * (b_distisoft - b_crf_2cut3i * r2 - crf_cut);
    double e_lj = lp->a_lj_lambda_n * a_e_lj + lp->b_lj_lambda_n * b_e_lj;
    double e_crf = (lp->a_crf_lambda_n * a_e_crf + lp->b_crf_lambda_n * b_e_crf) * fpepsi;

    double de_lj = -2.0 * alpha_lj *
                       (lp->a_lj_lambda_n * lp->b_ljs_lambda * a_c126 * a_dist6isoft * a_dist6isoft *
                            (2.0 * a_c12 * a_dist6isoft - a_c6) -
                        lp->b_lj_lambda_n * lp->a_ljs_lambda * b_c126 * b_dist6isoft * b_dist6isoft *
                            (2.0 * b_c12 * b_dist6isoft - b_c6)) +
                   (double)lp->lambda_exp * (lp->b_lj_lambda_n_1 * b_e_lj - lp->a_lj_lambda_n_1 * a_e_lj);
    double de_crf = -(lp->a_crf_lambda_n * a_q * lp->b_crfs_lambda * (a_dist3isoft - a_crf_pert * r2) -
                      lp->b_crf_lambda_n * b_q * lp->a_crfs_lambda * (b_dist3isoft - b_crf_pert * r2)) *
                        fpepsi * alpha_crf +
                    (double)lp->lambda_exp *
                        (lp->b_crf_lambda_n_1 * b_e_crf - lp->a_crf_lambda_n_1 * a_e_crf) * fpepsi;
    out[0] = force; out[1] = e_lj; out[2] = e_crf; out[3] = de_lj; out[4] = de_crf;
}

static void alphas(const or_pert *pt, int i, int j, double *al, double *ac) {
    int pi = pt->flag[i] != 0, pj = pt->flag[j] != 0;
    if (pi && pj) {
        *al = (pt->alpha_lj[i] + pt->alpha_lj[j]) * 0.5;
        *ac = (pt->alpha_crf[i] + pt->alpha_crf[j]) * 0.5;
    } else if (pi) {
        *al = pt->alpha_lj[i]; *ac = pt->alpha_crf[i];
    } else {
        *al = pt->alpha_lj[j]; *ac = pt->alpha_crf[j];
    }
}

/* perturbed_pairlist_correction (perturbed.rs:309-386): only pairs whose FIRST atom
 * is perturbed; state A from lj_crf_interaction, per-atom nearest image */
void or_perturbed_pairlist_correction(const double *pos, const uint32_t *pairs, int64_t np,
                                      const or_pert *pt, const double *q, const int32_t *iac,
                                      const double *lj, int32_t nt, const or_crf *crf,
                                      const or_lambda *lp, const or_pbc *pbc, double fpepsi,
                                      int n_atoms, or_pertcorr *out) {
    (void)n_atoms;
    for (int64_t k = 0; k < np; ++k) {
        int i = (int)pairs[2 * k], j = (int)pairs[2 * k + 1];
        if (!pt->flag[i]) continue;
        int pj = pt->flag[j] != 0;
        int a_iac_i = pt->a_iac[i], b_iac_i = pt->b_iac[i];
        int a_iac_j = pj ? pt->a_iac[j] : iac[j], b_iac_j = pj ? pt->b_iac[j] : iac[j];
        double a_q_i = pt->a_q[i], b_q_i = pt->b_q[i];
        double a_q_j = pj ? pt->a_q[j] : q[j], b_q_j = pj ? pt->b_q[j] : q[j];
        const double *la = lj + 4 * ((size_t)a_iac_i * nt + a_iac_j);
        const double *lb = lj + 4 * ((size_t)b_iac_i * nt + b_iac_j);
        double a_q = a_q_i * a_q_j, b_q = b_q_i * b_q_j;
        double al, ac;
        alphas(pt, i, j, &al, &ac);
        double r[3];
        or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
        double fa, elja, ecrfa, o[5];
        or_lj_crf_interaction(r, la[0], la[1], a_q * fpepsi, crf, &fa, &elja, &ecrfa);
        or_perturbed_lj_crf_interaction(r, la[0], la[1], lb[0], lb[1], a_q, b_q, al, ac, lp, crf, fpepsi, o);
        double d = o[0] - fa;
        for (int a = 0; a < 3; ++a) { double df = r[a] * d; out->f[3 * i + a] += df; out->f[3 * j + a] -= df; }
        out->de_lj += o[1] - elja;
        out->de_crf += o[2] - ecrfa;
        out->dhdl += o[3] + o[4];
    }
}

/* perturbed_self_energy_correction (perturbed.rs:395-425) */
void or_perturbed_self_energy_correction(const or_topo *t, const or_pert *pt, const or_crf *crf,
                                         const or_lambda *lp, double fpepsi, double *de,
                                         double *dhdl) {
    double e = 0.0, d = 0.0, n = (double)lp->lambda_exp;
    for (int i = 0; i < t->n_atoms; ++i) {
        if (!pt->flag[i]) continue;
        double qa2 = pt->a_q[i] * pt->a_q[i], qb2 = pt->b_q[i] * pt->b_q[i];
        double e_rf = (-(lp->a_crf_lambda_n * qa2) - lp->b_crf_lambda_n * qb2 + qa2) * crf->crf_cut * fpepsi;
        e += 0.5 * e_rf;
        double de_rf = n * (lp->a_crf_lambda_n_1 * qa2 - lp->b_crf_lambda_n_1 * qb2) * crf->crf_cut * fpepsi;
        d += 0.5 * de_rf;
    }
    *de = e;
    *dhdl = d;
}

/* perturbed_excluded_correction (perturbed.rs:431-511): soft cutoff scaled by
 * sqrt((rc^2/cut2soft)^3) as written at :474-475, no 1/r term */
void or_perturbed_excluded_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                      const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                      double fpepsi, or_pertcorr *out) {
    const double *q = t->charge;
    for (int i = 0; i < t->n_solute_atoms; ++i) {
        int pi = pt->flag[i] != 0;
        for (int k = t->excl_off[i]; k < t->excl_off[i + 1]; ++k) {
            int j = t->excl[k];
            if (j <= i) continue;
            int pj = pt->flag[j] != 0;
            if (!pi && !pj) continue;
            double a_q_i = pi ? pt->a_q[i] : q[i], b_q_i = pi ? pt->b_q[i] : q[i];
            double a_q_j = pj ? pt->a_q[j] : q[j], b_q_j = pj ? pt->b_q[j] : q[j];
            double al, ac;
            alphas(pt, i, j, &al, &ac);
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
            double r2 = len2(r);
            double a_q = a_q_i * a_q_j, b_q = b_q_i * b_q_j;
            double cut2 = crf->cutoff_sq;
            double a_cut2soft = cut2 + ac * lp->b_crfs_lambda2;
            double b_cut2soft = cut2 + ac * lp->a_crfs_lambda2;
            double a_crf_2cut3i = crf->crf_2cut3i * sqrt(powi_(cut2 / a_cut2soft, 3));
            double b_crf_2cut3i = crf->crf_2cut3i * sqrt(powi_(cut2 / b_cut2soft, 3));
            double a_crf_cut3i = 2.0 * a_crf_2cut3i;
            double b_crf_cut3i = 2.0 * b_crf_2cut3i;
            double a_crf_pert = 3.0 * a_crf_2cut3i / a_cut2soft;
            double b_crf_pert = 3.0 * b_crf_2cut3i / b_cut2soft;
            double a_e_crf = a_q * (-a_crf_2cut3i * r2 - crf->crf_cut);
            double b_e_crf = b_q * (-b_crf_2cut3i * r2 - crf->crf_cut);
            double e_crf_p = (lp->a_crf_lambda_n * a_e_crf + lp->b_crf_lambda_n * b_e_crf) * fpepsi;
            double f_p = (lp->a_crf_lambda_n * a_q * a_crf_cut3i + lp->b_crf_lambda_n * b_q * b_crf_cut3i) * fpepsi;
            double e_crf_a = a_q * fpepsi * (-crf->crf_2cut3i * r2 - crf->crf_cut);
            double f_a = a_q * fpepsi * crf->crf_cut3i;
            double n_exp = (double)lp->lambda_exp;
            double de_crf = ((lp->a_crf_lambda_n * a_q * lp->b_crfs_lambda * a_crf_pert -
                              lp->b_crf_lambda_n * b_q * lp->a_crfs_lambda * b_crf_pert) * r2 * ac +
                             n_exp * (lp->b_crf_lambda_n_1 * b_e_crf - lp->a_crf_lambda_n_1 * a_e_crf)) * fpepsi;
            double d = f_p - f_a;
            for (int a = 0; a < 3; ++a) { double df = r[a] * d; out->f[3 * i + a] += df; out->f[3 * j + a] -= df; }
            out->de_crf += e_crf_p - e_crf_a;
            out->dhdl += de_crf;
        }
    }
}

/* perturbed_one_four_correction (perturbed.rs:517-605) */
void or_perturbed_one_four_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                      const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                      double fpepsi, or_pertcorr *out) {
    if (!t->o14_off) return;
    const double *q = t->charge;
    const int32_t *iac = t->iac;
    int nt = t->n_types;
    for (int i = 0; i < t->n_atoms; ++i) {
        int pi = pt->flag[i] != 0;
        for (int k = t->o14_off[i]; k < t->o14_off[i + 1]; ++k) {
            int j = t->o14[k];
            if (j <= i) continue;
            int pj = pt->flag[j] != 0;
            if (!pi && !pj) continue;
            int a_iac_i = pi ? pt->a_iac[i] : iac[i], b_iac_i = pi ? pt->b_iac[i] : iac[i];
            int a_iac_j = pj ? pt->a_iac[j] : iac[j], b_iac_j = pj ? pt->b_iac[j] : iac[j];
            double a_q_i = pi ? pt->a_q[i] : q[i], b_q_i = pi ? pt->b_q[i] : q[i];
            double a_q_j = pj ? pt->a_q[j] : q[j], b_q_j = pj ? pt->b_q[j] : q[j];
            const double *la = t->lj + 4 * ((size_t)a_iac_i * nt + a_iac_j);
            const double *lb = t->lj + 4 * ((size_t)b_iac_i * nt + b_iac_j);
            double a_q = a_q_i * a_q_j, b_q = b_q_i * b_q_j;
            double al, ac;
            alphas(pt, i, j, &al, &ac);
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
            double r2 = len2(r);
            if (r2 < 1e-10) continue;
            double inv_r2 = 1.0 / r2;
            double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            double inv_r = sqrt(inv_r2);
            double e_lj_a = (la[3] * inv_r6 - la[2]) * inv_r6;
            double f_lj_a = (12.0 * la[3] * inv_r6 - 6.0 * la[2]) * inv_r6 * inv_r2;
            double e_crf_a = a_q * fpepsi * (inv_r - crf->crf_2cut3i * r2 - crf->crf_cut);
            double f_crf_a = a_q * fpepsi * (inv_r * inv_r2 + crf->crf_cut3i);
            double o[5];
            or_perturbed_lj_crf_interaction(r, la[2], la[3], lb[2], lb[3], a_q, b_q, al, ac, lp, crf, fpepsi, o);
            double d = o[0] - (f_lj_a + f_crf_a);
            for (int a = 0; a < 3; ++a) { double df = r[a] * d; out->f[3 * i + a] += df; out->f[3 * j + a] -= df; }
            out->de_lj += o[1] - e_lj_a;
            out->de_crf += o[2] - e_crf_a;
            out->dhdl += o[3] + o[4];
        }
    }
}

/* perturbed_atom_pair_correction (perturbed.rs:624-707): only pairs in the 1-4 lists */
void or_perturbed_atom_pair_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                       const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                       double fpepsi, or_pertcorr *out) {
    if (!t->o14_off) return;
    double n_exp = (double)lp->lambda_exp;
    for (int k = 0; k < pt->n_atom_pairs; ++k) {
        int i = pt->atom_pairs[4 * k], j = pt->atom_pairs[4 * k + 1];
        int at = pt->atom_pairs[4 * k + 2], bt = pt->atom_pairs[4 * k + 3];
        int in14 = contains_i32(t->o14 + t->o14_off[i], t->o14_off[i + 1] - t->o14_off[i], j) ||
                   contains_i32(t->o14 + t->o14_off[j], t->o14_off[j + 1] - t->o14_off[j], i);
        if (!in14) continue;
        double r[3];
        or_nearest_image(pbc, pos + 3 * i, pos + 3 * j, r);
        double r2 = len2(r);
        if (r2 < 1e-10) continue;
        double inv_r2 = 1.0 / r2;
        double inv_r6 = inv_r2 * inv_r2 * inv_r2;
        double inv_r = sqrt(inv_r2);
        const double *p = t->lj + 4 * ((size_t)t->iac[i] * t->n_types + t->iac[j]);
        double e_lj_reg = (p[3] * inv_r6 - p[2]) * inv_r6;
        double f_lj_reg = (12.0 * p[3] * inv_r6 - 6.0 * p[2]) * inv_r6 * inv_r2;
        double qp = t->charge[i] * t->charge[j] * fpepsi;
        double e_crf_reg = qp * (inv_r - crf->crf_2cut3i * r2 - crf->crf_cut);
        double f_crf_reg = qp * (inv_r * inv_r2 + crf->crf_cut3i);
        double e_lj_a, f_lj_a, e_lj_b = 0.0, f_lj_b = 0.0;
        if (at == 0) {
            e_lj_a = (p[1] * inv_r6 - p[0]) * inv_r6;
            f_lj_a = (12.0 * p[1] * inv_r6 - 6.0 * p[0]) * inv_r6 * inv_r2;
        } else {
            e_lj_a = (p[3] * inv_r6 - p[2]) * inv_r6;
            f_lj_a = (12.0 * p[3] * inv_r6 - 6.0 * p[2]) * inv_r6 * inv_r2;
        }
        double e_crf_a = e_crf_reg, f_crf_a = f_crf_reg, e_crf_b = 0.0, f_crf_b = 0.0;
        if (bt >= 0) {
            if (bt == 0) {
                e_lj_b = (p[1] * inv_r6 - p[0]) * inv_r6;
                f_lj_b = (12.0 * p[1] * inv_r6 - 6.0 * p[0]) * inv_r6 * inv_r2;
            } else {
                e_lj_b = (p[3] * inv_r6 - p[2]) * inv_r6;
                f_lj_b = (12.0 * p[3] * inv_r6 - 6.0 * p[2]) * inv_r6 * inv_r2;
            }
            e_crf_b = e_crf_reg;
            f_crf_b = f_crf_reg;
        }
        double e_lj_p = lp->a_lj_lambda_n * e_lj_a + lp->b_lj_lambda_n * e_lj_b;
        double e_crf_p = lp->a_crf_lambda_n * e_crf_a + lp->b_crf_lambda_n * e_crf_b;
        double f_lj_p = lp->a_lj_lambda_n * f_lj_a + lp->b_lj_lambda_n * f_lj_b;
        double f_crf_p = lp->a_crf_lambda_n * f_crf_a + lp->b_crf_lambda_n * f_crf_b;
        double de_lj = n_exp * (lp->b_lj_lambda_n_1 * e_lj_b - lp->a_lj_lambda_n_1 * e_lj_a);
        double de_crf = n_exp * (lp->b_crf_lambda_n_1 * e_crf_b - lp->a_crf_lambda_n_1 * e_crf_a);
        double d = (f_lj_p - f_lj_reg) + (f_crf_p - f_crf_reg);
        for (int a = 0; a < 3; ++a) { double df = r[a] * d; out->f[3 * i + a] += df; out->f[3 * j + a] -= df; }
        out->de_lj += e_lj_p - e_lj_reg;
        out->de_crf += e_crf_p - e_crf_reg;
        out->dhdl += de_lj + de_crf;
    }
}

/* ------------------------------------------------------------------ */
/* covalent terms (crates/gromos-forces/src/bonded), no periodic image */
/* (the reference takes pos(i) - pos(j) directly)                      */
/* ------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(const double *a, const double *b, double *r) { for (int k = 0; k < 3; ++k) r[k] = a[k] - b[k]; }
static inline void cross3(const double *a, const double *b, double *r) {   /* glam DVec3::cross */
    r[0] = a[1] * b[2] - b[1] * a[2];
    r[1] = a[2] * b[0] - b[2] * a[0];
    r[2] = a[0] * b[1] - b[0] * a[1];
}
/* glam DVec3::length: sqrt of the dot product */
static inline double length3(const double *a) { return sqrt(dot3(a, a)); }

/* calculate_bond_forces_quartic (bonds.rs:12-50) */
static double bonds_quartic(const or_bonded *b, const double *pos, double *f, double *vir) {
    double e = 0.0;
    for (int k = 0; k < b->nb; ++k) {
        int i = b->bonds[3 * k], j = b->bonds[3 * k + 1], t = b->bonds[3 * k + 2];
        if (t >= b->nbt) continue;
        const double *p = b->bond_types + 3 * t;
        double r[3];
        sub3(pos + 3 * i, pos + 3 * j, r);
        double r2 = dot3(r, r);
        double r0_2 = p[2] * p[2];
        double dr2 = r2 - r0_2;
        e += 0.25 * p[0] * dr2 * dr2;
        double fs = -p[0] * dr2, fv[3];
        for (int a = 0; a < 3; ++a) { fv[a] = r[a] * fs; f[3 * i + a] += fv[a]; f[3 * j + a] -= fv[a]; }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vir[3 * a + c] += r[a] * fv[c];
    }
    return e;
}

/* calculate_angle_forces, cosine-harmonic (angles.rs:12-75) */
static double angles_cosine(const or_bonded *b, const double *pos, double *f, double *vir) {
    double e = 0.0;
    for (int q = 0; q < b->na; ++q) {
        int i = b->angles[4 * q], j = b->angles[4 * q + 1], k = b->angles[4 * q + 2], t = b->angles[4 * q + 3];
        if (t >= b->nat) continue;
        const double *p = b->angle_types + 3 * t;
        double rij[3], rkj[3];
        sub3(pos + 3 * i, pos + 3 * j, rij);
        sub3(pos + 3 * k, pos + 3 * j, rkj);
        double dij = length3(rij), dkj = length3(rkj);
        if (dij < 1e-10 || dkj < 1e-10) continue;
        double cost = dot3(rij, rkj) / (dij * dkj);
        if (cost < -1.0) cost = -1.0;
        if (cost > 1.0) cost = 1.0;
        double cos0 = cos(p[2]);
        double d_cos = cost - cos0;
        e += 0.5 * p[0] * d_cos * d_cos;
        double df = -p[0] * d_cos;
        double fi[3], fk[3];
        for (int a = 0; a < 3; ++a) {
            fi[a] = (rkj[a] * (1.0 / dkj) - rij[a] * (cost / dij)) * (df / dij);
            fk[a] = (rij[a] * (1.0 / dij) - rkj[a] * (cost / dkj)) * (df / dkj);
        }
        for (int a = 0; a < 3; ++a) {
            double fj = -(fi[a] + fk[a]);
            f[3 * i + a] += fi[a];
            f[3 * j + a] += fj;
            f[3 * k + a] += fk[a];
        }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vir[3 * a + c] += rij[a] * fi[c] + rkj[a] * fk[c];
    }
    return e;
}

/* powi for the Chebyshev terms: squaring as LLVM expands a constant power */
static inline double pw(double x, int n) { return powi_(x, n); }

/* calculate_dihedral_forces (dihedrals.rs:19-130) */
static double dihedrals_cheb(const or_bonded *b, const double *pos, double *f, double *vir) {
    double e = 0.0;
    for (int q = 0; q < b->nd; ++q) {
        const int32_t *d = b->dihedrals + 5 * q;
        int i = d[0], j = d[1], k = d[2], l = d[3], t = d[4];
        if (t >= b->ndt) continue;
        const double *p = b->dihedral_types + 3 * t;
        double K = p[0], cospd = p[1];
        int m = (int)p[2];
        double rij[3], rkj[3], rkl[3];
        sub3(pos + 3 * i, pos + 3 * j, rij);
        sub3(pos + 3 * k, pos + 3 * j, rkj);
        sub3(pos + 3 * k, pos + 3 * l, rkl);
        double dkj2 = dot3(rkj, rkj);
        if (dkj2 < 1e-10) continue;
        double frim = dot3(rij, rkj) / dkj2, frln = dot3(rkl, rkj) / dkj2;
        double rim[3], rln[3];
        for (int a = 0; a < 3; ++a) {
            rim[a] = rij[a] - rkj[a] * frim;
            rln[a] = rkj[a] * frln - rkl[a];
        }
        double dim = length3(rim), dln = length3(rln);
        if (dim < 1e-10 || dln < 1e-10) continue;
        double c = dot3(rim, rln) / (dim * dln);
        if (c < -1.0) c = -1.0;
        if (c > 1.0) c = 1.0;
        double cm, dcm;
        switch (m) {
            case 0: cm = 0.0; dcm = 0.0; break;
            case 1: cm = c; dcm = 1.0; break;
            case 2: cm = 2.0 * c * c - 1.0; dcm = 4.0 * c; break;
            case 3: cm = 4.0 * pw(c, 3) - 3.0 * c; dcm = 12.0 * c * c - 3.0; break;
            case 4: cm = 8.0 * pw(c, 4) - 8.0 * c * c + 1.0; dcm = 32.0 * pw(c, 3) - 16.0 * c; break;
            case 5: cm = 16.0 * pw(c, 5) - 20.0 * pw(c, 3) + 5.0 * c; dcm = 80.0 * pw(c, 4) - 60.0 * c * c + 5.0; break;
            case 6:
                cm = 32.0 * pw(c, 6) - 48.0 * pw(c, 4) + 18.0 * c * c - 1.0;
                dcm = 192.0 * pw(c, 5) - 192.0 * pw(c, 3) + 36.0 * c;
                break;
            default: continue;   /* unsupported multiplicity: the term is skipped */
        }
        e += K * (1.0 + cospd * cm);
        double ki = -K * cospd * dcm / dim, kl = -K * cospd * dcm / dln;
        double kj1 = frim - 1.0, kj2 = frln;
        double fi[3], fl[3], fj[3], fk[3], rlj[3];
        for (int a = 0; a < 3; ++a) {
            fi[a] = rln[a] * (ki / dln) - rim[a] * (ki * c / dim);
            fl[a] = rim[a] * (kl / dim) - rln[a] * (kl * c / dln);
            fj[a] = fi[a] * kj1 - fl[a] * kj2;
            fk[a] = -(fi[a] + fj[a] + fl[a]);
            f[3 * i + a] += fi[a];
            f[3 * j + a] += fj[a];
            f[3 * k + a] += fk[a];
            f[3 * l + a] += fl[a];
        }
        sub3(pos + 3 * l, pos + 3 * j, rlj);
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) vir[3 * a + bb] += rij[a] * fi[bb] + rkj[a] * fk[bb] + rlj[a] * fl[bb];
    }
    return e;
}

/* calculate_improper_dihedral_forces (improper.rs:12-105) */
static double impropers_harm(const or_bonded *b, const double *pos, double *f, double *vir) {
    const double PI = 3.14159265358979323846;
    double e = 0.0;
    for (int q = 0; q < b->ni; ++q) {
        const int32_t *d = b->impropers + 5 * q;
        int i = d[0], j = d[1], k = d[2], l = d[3], t = d[4];
        if (t >= b->nit) continue;
        const double *p = b->improper_types + 2 * t;
        double K = p[0], q0 = p[1];
        double rkj[3], rij[3], rkl[3], rmj[3], rnk[3];
        sub3(pos + 3 * k, pos + 3 * j, rkj);
        sub3(pos + 3 * i, pos + 3 * j, rij);
        sub3(pos + 3 * k, pos + 3 * l, rkl);
        cross3(rij, rkj, rmj);
        cross3(rkj, rkl, rnk);
        double dkj2 = dot3(rkj, rkj), dmj2 = dot3(rmj, rmj), dnk2 = dot3(rnk, rnk);
        double dkj = sqrt(dkj2), dmj = sqrt(dmj2), dnk = sqrt(dnk2);
        if (dmj < 1e-10 || dnk < 1e-10) continue;
        double acs = dot3(rmj, rnk) / (dmj * dnk);
        if (acs < -1.0) acs = -1.0;
        if (acs > 1.0) acs = 1.0;
        double zeta = acos(acs);
        if (dot3(rij, rnk) < 0.0) zeta = -zeta;
        double za = zeta;
        while (za < (q0 - PI)) za += 2.0 * PI;
        while (za > (q0 + PI)) za -= 2.0 * PI;
        double dz = za - q0;
        e += 0.5 * K * dz * dz;
        double ki = -K * dz * dkj, kl = -ki;
        if (dmj2 < 1.0e-10 * dkj2) ki = 0.0; else ki /= dmj2;
        if (dnk2 < 1.0e-10 * dkj2) kl = 0.0; else kl /= dnk2;
        double kj1 = dot3(rij, rkj) / dkj2 - 1.0, kj2 = dot3(rkl, rkj) / dkj2;
        double fi[3], fl[3], fj[3], fk[3], rlj[3];
        for (int a = 0; a < 3; ++a) {
            fi[a] = rmj[a] * ki;
            fl[a] = rnk[a] * kl;
            fj[a] = fi[a] * kj1 - fl[a] * kj2;
            fk[a] = -(fi[a] + fj[a] + fl[a]);
            f[3 * i + a] += fi[a];
            f[3 * j + a] += fj[a];
            f[3 * k + a] += fk[a];
            f[3 * l + a] += fl[a];
        }
        sub3(pos + 3 * l, pos + 3 * j, rlj);
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) vir[3 * a + bb] += rij[a] * fi[bb] + rkj[a] * fk[bb] + rlj[a] * fl[bb];
    }
    return e;
}

void or_bonded_forces(const or_bonded *b, int n, const double *pos, double *f, double vir[9], or_bonded_e *e) {
    size_t nb = sizeof(double) * 3 * (size_t)n;
    double *part = (double *)malloc(nb);
    double pv[9];
    memset(f, 0, nb);
    for (int c = 0; c < 9; ++c) vir[c] = 0.0;
    double en[4];
    /* NTF gates (md.rs:331-335): bonds ntf[0], angles ntf[1], impropers ntf[2], dihedrals ntf[3] */
    const int on[4] = {b->ntf[0], b->ntf[1], b->ntf[3], b->ntf[2]};
    for (int kind = 0; kind < 4; ++kind) {
        en[kind] = 0.0;
        if (!on[kind]) continue;
        memset(part, 0, nb);
        for (int c = 0; c < 9; ++c) pv[c] = 0.0;
        switch (kind) {
            case 0: en[0] = bonds_quartic(b, pos, part, pv); break;
            case 1: en[1] = angles_cosine(b, pos, part, pv); break;
            case 2: en[2] = dihedrals_cheb(b, pos, part, pv); break;
            default: en[3] = impropers_harm(b, pos, part, pv); break;
        }
        /* ForceEnergy::add (bonded/mod.rs:54-62) */
        for (size_t k = 0; k < 3 * (size_t)n; ++k) f[k] += part[k];
        for (int c = 0; c < 9; ++c) vir[c] += pv[c];
    }
    free(part);
    e->e_bond = en[0];
    e->e_angle = en[1];
    e->e_dihedral = en[2];
    e->e_improper = en[3];
}

/* ------------------------------------------------------------------ */
/* Forcefield::apply, nonbonded part (forcefield.rs:226-1060)          */
/* ------------------------------------------------------------------ */
struct or_ff {
    const or_topo *t;
    or_pbc pbc;
    or_ff_params prm;
    or_crf crf;
    or_pert pt;
    int has_pert;
    or_pairlist pl;
    int steps_since_update;
    int twin_range_active;
    int longrange_computed;
    int use_cg_grouped;
    uint32_t *groups; int64_t ngroups, groups_cap;
    int cached;
    or_storage nb, lr;
    int64_t last_pairs;
    double *bf;                  /* covalent forces of the last apply */
    or_bonded_e be;
};

or_ff *or_ff_create(const or_topo *t, const or_pbc *pbc, const or_ff_params *prm, const or_pert *pt) {
    or_ff *ff = (or_ff *)calloc(1, sizeof(or_ff));
    ff->t = t;
    ff->pbc = *pbc;
    ff->prm = *prm;
    or_crf_new(prm->rcutl, 1.0, prm->epsrf, prm->kappa, &ff->crf);
    ff->twin_range_active = prm->rcutp < prm->rcutl - 1e-10;   /* forcefield.rs:126 */
    if (pt && pt->flag) {
        ff->pt = *pt;
        for (int i = 0; i < t->n_atoms; ++i) if (pt->flag[i]) { ff->has_pert = 1; break; }
    }
    or_storage_init(&ff->nb, t->n_atoms);
    or_storage_init(&ff->lr, t->n_atoms);
    if (t->bonded) ff->bf = (double *)calloc(3 * (size_t)(t->n_atoms > 0 ? t->n_atoms : 1), sizeof(double));
    /* use_cg_grouped_solute (forcefield.rs:271-283) */
    int cg_avail = t->n_cg > 0;
    if (pbc->kind == 0) ff->use_cg_grouped = cg_avail;
    else if (pbc->kind == 2) ff->use_cg_grouped = 0;   /* "be conservative for triclinic" */
    else {
        double hm = pbc->half[0];
        if (pbc->half[1] < hm) hm = pbc->half[1];
        if (pbc->half[2] < hm) hm = pbc->half[2];
        ff->use_cg_grouped = cg_avail && (hm > prm->rcutl + prm->skin + 0.3);
    }
    return ff;
}

void or_ff_destroy(or_ff *ff) {
    if (!ff) return;
    or_pairlist_free(&ff->pl);
    or_storage_free(&ff->nb);
    or_storage_free(&ff->lr);
    free(ff->groups);
    free(ff->bf);
    free(ff);
}

static void ff_update_pairlist(or_ff *ff, const double *pos) {
    if (ff->prm.algorithm == 2)
        or_pairlist_cell(ff->t, pos, &ff->pbc, ff->prm.rcutp, ff->prm.rcutl, ff->prm.skin, &ff->pl);
    else
        or_pairlist_standard(ff->t, pos, &ff->pbc, ff->prm.rcutp, ff->prm.rcutl, ff->prm.skin, &ff->pl);
    ff->steps_since_update = 0;
}

void or_ff_initial_pairlist(or_ff *ff, const double *pos) { ff_update_pairlist(ff, pos); }
const or_pairlist *or_ff_pairlist(const or_ff *ff) { return &ff->pl; }
void or_ff_last_bonded(const or_ff *ff, or_bonded_e *e) { *e = ff->be; }
int or_ff_used_cg_grouped(const or_ff *ff) { return ff->use_cg_grouped; }
int64_t or_ff_last_pair_count(const or_ff *ff) { return ff->last_pairs; }

/* CG-pair runs of solute_short (forcefield.rs:321-358) */
static void build_groups(or_ff *ff) {
    const or_topo *t = ff->t;
    ff->ngroups = 0;
    const or_plist *l = &ff->pl.solute_short;
    if (l->n == 0) return;
    int *a2cg = (int *)malloc(sizeof(int) * t->n_atoms);
    for (int cg = 0; cg < t->n_cg; ++cg)
        for (int a = t->cg_start[cg]; a < t->cg_start[cg + 1]; ++a) a2cg[a] = cg;
    uint32_t start = 0;
    int ci = a2cg[l->p[0]], cj = a2cg[l->p[1]];
    for (int64_t k = 1; k <= l->n; ++k) {
        int gi = -1, gj = -1;
        if (k < l->n) { gi = a2cg[l->p[2 * k]]; gj = a2cg[l->p[2 * k + 1]]; }
        if (k == l->n || gi != ci || gj != cj) {
            if (ff->ngroups == ff->groups_cap) {
                ff->groups_cap = ff->groups_cap ? 2 * ff->groups_cap : 1024;
                ff->groups = (uint32_t *)realloc(ff->groups, sizeof(uint32_t) * 4 * (size_t)ff->groups_cap);
            }
            uint32_t *G = ff->groups + 4 * ff->ngroups++;
            G[0] = (uint32_t)t->cg_start[ci];
            G[1] = (uint32_t)t->cg_start[cj];
            G[2] = start;
            G[3] = (uint32_t)k;
            start = (uint32_t)k;
            ci = gi; cj = gj;
        }
    }
    free(a2cg);
}

int or_ff_apply(or_ff *ff, const double *pos, double *forces, double *e_lj, double *e_crf,
                double *vir, double *dhdl) {
    const or_topo *t = ff->t;
    int n = t->n_atoms;
    const double *q = t->charge;
    const int32_t *iac = t->iac;
    int nt = t->n_types;
    double fpe = ff->prm.fpepsi;
    int par = ff->prm.parallel, virial = ff->prm.virial;
    int64_t npairs = 0;

    int updated = ff->steps_since_update >= ff->prm.nsnb;            /* pairlist.rs:105-107 */
    if (updated) ff_update_pairlist(ff, pos);
    ff->steps_since_update++;
    if (updated || ff->pl.solute_short.n == 0 || !ff->cached) {
        if (ff->use_cg_grouped) build_groups(ff); else ff->ngroups = 0;
        ff->cached = 1;
    }

    or_storage *nb = &ff->nb;
    or_storage_clear(nb);
    const or_plist *ss = &ff->pl.solute_short;
    /* solute short: CG-grouped or per-atom (forcefield.rs:420-535) */
    if (ff->use_cg_grouped) {
        if (par) {
            or_storage tmp; or_storage_init(&tmp, n);
            or_cg_grouped_parallel(pos, q, iac, ss->p, ff->groups, ff->ngroups, t->lj, nt, &ff->crf, &ff->pbc, fpe, virial, &tmp);
            or_storage_merge(nb, &tmp); or_storage_free(&tmp);
        } else {
            or_lj_crf_innerloop_cg_grouped(pos, q, iac, ss->p, ff->groups, ff->ngroups, t->lj, nt, &ff->crf, &ff->pbc, fpe, virial, nb);
        }
    } else {
        if (par) {
            or_storage tmp; or_storage_init(&tmp, n);
            or_lj_crf_innerloop_parallel(pos, q, iac, ss->p, ss->n, t->lj, nt, &ff->crf, &ff->pbc, fpe, virial, &tmp);
            or_storage_merge(nb, &tmp); or_storage_free(&tmp);
        } else {
            or_lj_crf_innerloop(pos, q, iac, ss->p, ss->n, t->lj, nt, &ff->crf, &ff->pbc, fpe, virial, nb);
        }
    }
    npairs += ss->n;
    /* solvent short (forcefield.rs:544-608) */
    const or_plist *vs = &ff->pl.solvent_short;
    int aps = t->atoms_per_solvent;
    if (vs->n > 0) {
        if (par) {
            or_storage tmp; or_storage_init(&tmp, n);
            or_solvent_innerloop_parallel(pos, q, iac, vs->p, vs->n, t->lj, nt, &ff->crf, &ff->pbc, aps, fpe, virial, &tmp);
            or_storage_merge(nb, &tmp); or_storage_free(&tmp);
        } else {
            or_solvent_innerloop(pos, q, iac, vs->p, vs->n, t->lj, nt, &ff->crf, &ff->pbc, aps, fpe, virial, nb);
        }
        npairs += vs->n * (int64_t)aps * aps;
    }
    /* twin range (forcefield.rs:620-697) */
    const or_plist *sl = &ff->pl.solute_long, *vl = &ff->pl.solvent_long;
    if (ff->twin_range_active) {
        if (updated || !ff->longrange_computed) {
            or_storage *lr = &ff->lr;
            or_storage_clear(lr);
            if (sl->n) or_lj_crf_innerloop(pos, q, iac, sl->p, sl->n, t->lj, nt, &ff->crf, &ff->pbc, fpe, 1, lr);
            if (vl->n) or_lj_crf_innerloop(pos, q, iac, vl->p, vl->n, t->lj, nt, &ff->crf, &ff->pbc, fpe, 1, lr);
            ff->longrange_computed = 1;
            npairs += sl->n + vl->n;
        }
        or_storage_merge(nb, &ff->lr);
    } else {
        if (sl->n) { or_lj_crf_innerloop(pos, q, iac, sl->p, sl->n, t->lj, nt, &ff->crf, &ff->pbc, fpe, 1, nb); npairs += sl->n; }
        if (vl->n) { or_solvent_innerloop(pos, q, iac, vl->p, vl->n, t->lj, nt, &ff->crf, &ff->pbc, aps, fpe, 1, nb); npairs += vl->n * (int64_t)aps * aps; }
    }
    /* RF self + excluded (forcefield.rs:730-739), 1-4 (forcefield.rs:743-756) */
    or_rf_excluded_interactions(t, pos, &ff->crf, &ff->pbc, fpe, nb);
    or_one_four_interaction_loop(t, pos, &ff->crf, &ff->pbc, fpe, 1.0, nb);

    double dh = 0.0;
    if (ff->has_pert) {
        /* forcefield.rs:764-951 */
        or_lambda lp;
        double l = ff->prm.lambda;
        or_lambda_from(l, l, l, l, 1.0, 1.0, 1.0, 1.0, ff->prm.nlam, &lp);
        or_pertcorr c;
        c.f = (double *)calloc((size_t)3 * n, sizeof(double));
        const or_plist *lists[4] = {ss, sl, vs, vl};
        for (int li = 0; li < 4; ++li) {
            if (lists[li]->n == 0) continue;
            memset(c.f, 0, sizeof(double) * 3 * (size_t)n);
            c.de_lj = c.de_crf = c.dhdl = 0.0;
            or_perturbed_pairlist_correction(pos, lists[li]->p, lists[li]->n, &ff->pt, q, iac, t->lj, nt, &ff->crf, &lp, &ff->pbc, fpe, n, &c);
            nb->e_lj += c.de_lj; nb->e_crf += c.de_crf; dh += c.dhdl;
            for (int64_t k = 0; k < 3 * (int64_t)n; ++k) nb->f[k] += c.f[k];
        }
        double dse, dsd;
        or_perturbed_self_energy_correction(t, &ff->pt, &ff->crf, &lp, fpe, &dse, &dsd);
        nb->e_crf += dse; dh += dsd;
        memset(c.f, 0, sizeof(double) * 3 * (size_t)n); c.de_lj = c.de_crf = c.dhdl = 0.0;
        or_perturbed_excluded_correction(t, pos, &ff->pt, &ff->crf, &lp, &ff->pbc, fpe, &c);
        nb->e_crf += c.de_crf; dh += c.dhdl;
        for (int64_t k = 0; k < 3 * (int64_t)n; ++k) nb->f[k] += c.f[k];
        if (t->o14_off) {
            memset(c.f, 0, sizeof(double) * 3 * (size_t)n); c.de_lj = c.de_crf = c.dhdl = 0.0;
            or_perturbed_one_four_correction(t, pos, &ff->pt, &ff->crf, &lp, &ff->pbc, fpe, &c);
            nb->e_lj += c.de_lj; nb->e_crf += c.de_crf; dh += c.dhdl;
            for (int64_t k = 0; k < 3 * (int64_t)n; ++k) nb->f[k] += c.f[k];
        }
        if (ff->pt.n_atom_pairs > 0) {
            memset(c.f, 0, sizeof(double) * 3 * (size_t)n); c.de_lj = c.de_crf = c.dhdl = 0.0;
            or_perturbed_atom_pair_correction(t, pos, &ff->pt, &ff->crf, &lp, &ff->pbc, fpe, &c);
            nb->e_lj += c.de_lj; dh += c.dhdl;   /* forcefield.rs:923-928: only e_lj is added */
            for (int64_t k = 0; k < 3 * (int64_t)n; ++k) nb->f[k] += c.f[k];
        }
        free(c.f);
    }
    /* assembly (forcefield.rs:953-1000): force = bonded + nonbonded, virial nonbonded + bonded */
    if (t->bonded) {
        double bvir[9];
        or_bonded_forces(t->bonded, n, pos, ff->bf, bvir, &ff->be);
        if (forces)
            for (int64_t k = 0; k < 3 * (int64_t)n; ++k) forces[k] = ff->bf[k] + nb->f[k];
        if (vir)
            for (int c = 0; c < 9; ++c) vir[c] = nb->vir[c] + bvir[c];
    } else {
        memset(&ff->be, 0, sizeof ff->be);
        if (forces) memcpy(forces, nb->f, sizeof(double) * 3 * (size_t)n);
        if (vir) memcpy(vir, nb->vir, sizeof(double) * 9);
    }
    if (e_lj) *e_lj = nb->e_lj;
    if (e_crf) *e_crf = nb->e_crf;
    if (dhdl) *dhdl = dh;
    ff->last_pairs = npairs;
    return 0;
}

/* ------------------------------------------------------------------ */
/* step pieces                                                         */
/* ------------------------------------------------------------------ */
/* LeapFrogVelocity (leap_frog.rs:37-68) then LeapFrogPosition (leap_frog.rs:116-140) */
void or_leapfrog(int n, const double *im, double dt, const double *xo, const double *vo,
                 const double *fo, double *xn, double *vn) {
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            double acc = fo[3 * i + a] * im[i];
            vn[3 * i + a] = vo[3 * i + a] + acc * dt;
        }
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) xn[3 * i + a] = xo[3 * i + a] + vn[3 * i + a] * dt;
}

typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* glam DVec3::cross */
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline v3 ld(const double *p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st(double *p, int i, v3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

/* settle (constraints.rs:815-1036) */
int or_settle(int n_solute, int n_atoms, double m_o, double m_h, double dist_oh, double dist_hh,
              double dt, const double *pos_old, double *pos_new, double *vel_new,
              double *cons_force, double *vir) {
    double half_mo_div_mh = 0.5 * m_o / m_h;
    double rc = 0.5 * dist_hh;
    double ra = sqrt(dist_oh * dist_oh - rc * rc) / (1.0 + half_mo_div_mh);
    double rb = half_mo_div_mh * ra;
    double dt_i = 1.0 / dt;
    double dt2_i = dt_i * dt_i;
    int errors = 0;
    for (int i = n_solute; i + 3 <= n_atoms; i += 3) {
        v3 po0 = ld(pos_old, i), po1 = ld(pos_old, i + 1), po2 = ld(pos_old, i + 2);
        v3 pn0 = ld(pos_new, i), pn1 = ld(pos_new, i + 1), pn2 = ld(pos_new, i + 2);
        v3 b0 = vsub(po1, po0), c0 = vsub(po2, po0);
        v3 d0 = vdiv(vadd(vadd(vmul(pn0, m_o), vmul(pn1, m_h)), vmul(pn2, m_h)), m_o + m_h + m_h);
        v3 a1 = vsub(pn0, d0), b1 = vsub(pn1, d0), c1 = vsub(pn2, d0);
        v3 n0 = vcross(b0, c0);
        v3 n1 = vcross(a1, n0);
        v3 n2 = vcross(n0, n1);
        n0 = vdiv(n0, sqrt(vdot(n0, n0)));
        n1 = vdiv(n1, sqrt(vdot(n1, n1)));
        n2 = vdiv(n2, sqrt(vdot(n2, n2)));
        v3 m1 = mk(n1.x, n2.x, n0.x), m2 = mk(n1.y, n2.y, n0.y), m0 = mk(n1.z, n2.z, n0.z);
        b0 = mk(vdot(n1, b0), vdot(n2, b0), vdot(n0, b0));
        c0 = mk(vdot(n1, c0), vdot(n2, c0), vdot(n0, c0));
        a1 = mk(vdot(n1, a1), vdot(n2, a1), vdot(n0, a1));
        b1 = mk(vdot(n1, b1), vdot(n2, b1), vdot(n0, b1));
        c1 = mk(vdot(n1, c1), vdot(n2, c1), vdot(n0, c1));
        double sinphi = a1.z / ra;
        double omsp2 = 1.0 - sinphi * sinphi;
        if (omsp2 < 0.0) { errors++; continue; }
        double cosphi = sqrt(omsp2);
        double sinpsi = (b1.z - c1.z) / (2.0 * rc * cosphi);
        double omss2 = 1.0 - sinpsi * sinpsi;
        if (omss2 < 0.0) { errors++; continue; }
        double cospsi = sqrt(omss2);
        double minus_rb_cosphi = -rb * cosphi;
        double rc_cospsi = rc * cospsi;
        double rc_sinpsi_sinphi = rc * sinpsi * sinphi;
        double rc_sinpsi_cosphi = rc * sinpsi * cosphi;
        v3 a2 = mk(0.0, ra * cosphi, ra * sinphi);
        v3 b2 = mk(-rc_cospsi, minus_rb_cosphi - rc_sinpsi_sinphi, -rb * sinphi + rc_sinpsi_cosphi);
        v3 c2 = mk(rc_cospsi, minus_rb_cosphi + rc_sinpsi_sinphi, -rb * sinphi - rc_sinpsi_cosphi);
        double alpha = b2.x * (b0.x - c0.x) + b0.y * b2.y + c0.y * c2.y;
        double beta = b2.x * (c0.y - b0.y) + b0.x * b2.y + c0.x * c2.y;
        double gamma = b0.x * b1.y - b1.x * b0.y + c0.x * c1.y - c1.x * c0.y;
        double a2b2 = alpha * alpha + beta * beta;
        double sintheta = (alpha * gamma - beta * sqrt(a2b2 - gamma * gamma)) / a2b2;
        double omst2 = 1.0 - sintheta * sintheta;
        if (omst2 < 0.0) { errors++; continue; }
        double costheta = sqrt(omst2);
        v3 a3 = mk(-a2.y * sintheta, a2.y * costheta, a1.z);
        v3 b3 = mk(b2.x * costheta - b2.y * sintheta, b2.x * sintheta + b2.y * costheta, b1.z);
        v3 c3 = mk(-b2.x * costheta - c2.y * sintheta, -b2.x * sintheta + c2.y * costheta, c1.z);
        v3 pa = vadd(mk(vdot(a3, m1), vdot(a3, m2), vdot(a3, m0)), d0);
        v3 pb = vadd(mk(vdot(b3, m1), vdot(b3, m2), vdot(b3, m0)), d0);
        v3 pc = vadd(mk(vdot(c3, m1), vdot(c3, m2), vdot(c3, m0)), d0);
        v3 da = vsub(pa, pn0), db = vsub(pb, pn1), dc = vsub(pc, pn2);
        st(pos_new, i, pa); st(pos_new, i + 1, pb); st(pos_new, i + 2, pc);
        v3 cf[3] = {vmul(vmul(da, m_o), dt2_i), vmul(vmul(db, m_h), dt2_i), vmul(vmul(dc, m_h), dt2_i)};
        if (cons_force) { st(cons_force, i, cf[0]); st(cons_force, i + 1, cf[1]); st(cons_force, i + 2, cf[2]); }
        st(vel_new, i, vadd(ld(vel_new, i), vmul(da, dt_i)));
        st(vel_new, i + 1, vadd(ld(vel_new, i + 1), vmul(db, dt_i)));
        st(vel_new, i + 2, vadd(ld(vel_new, i + 2), vmul(dc, dt_i)));
        if (vir) {
            v3 p[3] = {po0, po1, po2};
            double dv[9] = {0};
            for (int k = 0; k < 3; ++k) {
                double pk[3] = {p[k].x, p[k].y, p[k].z}, fk[3] = {cf[k].x, cf[k].y, cf[k].z};
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) dv[3 * r + c] += pk[r] * fk[c];
            }
            for (int k = 0; k < 9; ++k) vir[k] += dv[k];
        }
    }
    return errors;
}

/* TemperatureCalculation (temperature.rs:63-87) */
double or_ekin(int n, const double *mass, const double *vn, const double *vo) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        double a = vn[3 * i] * vn[3 * i] + vn[3 * i + 1] * vn[3 * i + 1] + vn[3 * i + 2] * vn[3 * i + 2];
        double b = vo[3 * i] * vo[3 * i] + vo[3 * i + 1] * vo[3 * i + 1] + vo[3 * i + 2] * vo[3 * i + 2];
        e += 0.5 * mass[i] * (a + b) / 2.0;
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* MD driver (md.rs:1143-1286 sequence; configuration.rs double buffer) */
/* ------------------------------------------------------------------ */
struct or_md {
    const or_topo *t;
    or_ff *ff;
    double dt;
    int use_settle;
    int thermo;                 /* one bath over all atoms: 1 Berendsen, 2 Nose-Hoover (chain) */
    double t_ref, tau_t, dof_t;
    int nhc;                    /* Nose-Hoover chain length */
    double zeta[OR_NHC_MAX];    /* the chain's state, zero at start */
    int nticom, nscm;           /* centre-of-mass motion removal (INITIALISE NTICOM, COMTRANSROT NSCM) */
    int n_shake;                /* SHAKE constraints: the solute's, then the solvent's */
    int32_t *sh_i, *sh_j;
    double *sh_len;
    double sh_tol;
    int lc_order_u, lc_order_v; /* LINCS expansion orders: solute (NTCP0), solvent (NTCS0) */
    int lc_nu;                  /* the solute's LINCS constraints */
    int32_t *lc_ij;
    double *lc_len;
    int lc_nt;                  /* one solvent molecule's LINCS template */
    int32_t lc_tij[2 * OR_LINCS_TPL_MAX];
    double lc_tlen[OR_LINCS_TPL_MAX];
    int baro;                   /* Berendsen barostat, isotropic */
    double p_ref, comp, tau_p;
    int vtype;                  /* virial: 0 none, 1 atomic, 2 molecular */
    int n_pg;                   /* pressure groups: atoms [pg_start[g], pg_start[g + 1]) */
    int32_t *pg_start;
    double pressure;            /* the last step's isotropic pressure */
    int64_t step;               /* md steps done (SimulationState::step) */
    double *pos[2], *vel[2], *force[2];
    double *inv_mass;
    int cur;
    double m_o, m_h, d_oh, d_hh;
    int64_t pairs;
};

or_md *or_md_create(const or_topo *t, const or_pbc *pbc, const or_ff_params *prm,
                    const or_pert *pt, double dt, int use_settle, const double *pos,
                    const double *vel) {
    or_md *md = (or_md *)calloc(1, sizeof(or_md));
    int n = t->n_atoms;
    md->t = t;
    md->ff = or_ff_create(t, pbc, prm, pt);
    md->dt = dt;
    md->use_settle = use_settle;
    for (int k = 0; k < 2; ++k) {
        md->pos[k] = (double *)calloc((size_t)3 * n, sizeof(double));
        md->vel[k] = (double *)calloc((size_t)3 * n, sizeof(double));
        md->force[k] = (double *)calloc((size_t)3 * n, sizeof(double));
    }
    memcpy(md->pos[0], pos, sizeof(double) * 3 * (size_t)n);
    memcpy(md->vel[0], vel, sizeof(double) * 3 * (size_t)n);
    md->inv_mass = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) md->inv_mass[i] = t->mass[i] > 0.0 ? 1.0 / t->mass[i] : 0.0;
    if (use_settle && t->n_atoms > t->n_solute_atoms) {
        md->m_o = t->mass[t->n_solute_atoms];
        md->m_h = t->mass[t->n_solute_atoms + 1];
    }
    or_ff_initial_pairlist(md->ff, md->pos[0]);
    return md;
}

void or_md_set_settle_geometry(or_md *md, double d_oh, double d_hh);
void or_md_set_settle_geometry(or_md *md, double d_oh, double d_hh) { md->d_oh = d_oh; md->d_hh = d_hh; }

void or_md_destroy(or_md *md) {
    if (!md) return;
    free(md->pg_start);
    free(md->sh_i);
    free(md->sh_j);
    free(md->sh_len);
    free(md->lc_ij);
    free(md->lc_len);
    or_ff_destroy(md->ff);
    for (int k = 0; k < 2; ++k) { free(md->pos[k]); free(md->vel[k]); free(md->force[k]); }
    free(md->inv_mass);
    free(md);
}

void or_md_set_berendsen(or_md *md, double t_ref, double tau, double dof) {
    md->thermo = 1;
    md->t_ref = t_ref;
    md->tau_t = tau;
    md->dof_t = dof;
}

/* NoseHooverThermostat::new_single_bath / new_chain_bath
 * (nosehoover_thermostat.rs:64-113): the chain starts at zero */
int or_md_set_nosehoover(or_md *md, double t_ref, double tau, double dof, int nhc) {
    if (nhc < 1 || nhc > OR_NHC_MAX) return -1;
    md->thermo = 2;
    md->t_ref = t_ref;
    md->tau_t = tau;
    md->dof_t = dof;
    md->nhc = nhc;
    for (int i = 0; i < OR_NHC_MAX; ++i) md->zeta[i] = 0.0;
    return 0;
}

/* NoseHooverThermostat::apply (nosehoover_thermostat.rs:116-227), one bath:
 * advances the chain with E_kin_new (the velocities entering the step) and
 * returns the velocity scale; 1 where the reference leaves the velocities
 * untouched (no coupling, or |scale - 1| < 1e-15) */
double or_nosehoover_step(double *zeta, int nhc, double ekin, double t_ref, double tau, double dof, double dt) {
    if (tau < 0.0) return 1.0;
    double free_temp = dof > 0.0 ? 2.0 * ekin / (dof * OR_KB) : 0.0;
    if (free_temp < DBL_EPSILON) free_temp = t_ref;
    double scale;
    if (nhc == 1) {
        zeta[0] += dt / (tau * tau) * (free_temp / t_ref - 1.0);
        scale = 1.0 - zeta[0] * dt;
    } else {
        /* tail to head; tau_vec[0] = tau^2, tau_vec[i > 0] = tau^2 / dof */
        const double tau2 = tau * tau;
#define OR_TV(i) ((i) == 0 ? tau2 : tau2 / dof)
        {
            const double z = zeta[nhc - 2];
            zeta[nhc - 1] += (OR_TV(nhc - 2) * z * z - 1.0 / dof) / OR_TV(nhc - 1) * dt;
        }
        for (int i = nhc - 2; i >= 1; --i) {
            const double zp = zeta[i - 1], zc = zeta[i], zn = zeta[i + 1];
            zeta[i] += ((OR_TV(i - 1) * zp * zp - 1.0 / dof) / OR_TV(i) - zc * zn) * dt;
        }
        {
            const double z0 = zeta[0], z1 = zeta[1];
            zeta[0] += ((free_temp / t_ref - 1.0) / OR_TV(0) - z1 * z0) * dt;
        }
#undef OR_TV
        scale = 1.0 - zeta[0] * dt;
    }
    if (fabs(scale - 1.0) < 1e-15) return 1.0;
    return scale;
}

void or_md_get_nosehoover(const or_md *md, double *zeta) {
    for (int i = 0; i < md->nhc; ++i) zeta[i] = md->zeta[i];
}

/* The periodic box changed (the barostat): Forcefield::apply's per-step box
 * refresh (forcefield.rs:243-265: Rectangular::new, half box and reciprocal)
 * and its CG-grouped predicate (forcefield.rs:267-283), rectangular boxes */
void or_ff_set_box(or_ff *ff, const double L[3]) {
    for (int a = 0; a < 3; ++a) {
        ff->pbc.L[a] = L[a];
        ff->pbc.invL[a] = 1.0 / L[a];
        ff->pbc.half[a] = L[a] * 0.5;
    }
    double hm = ff->pbc.half[0];
    if (ff->pbc.half[1] < hm) hm = ff->pbc.half[1];
    if (ff->pbc.half[2] < hm) hm = ff->pbc.half[2];
    ff->use_cg_grouped = ff->t->n_cg > 0 && (hm > ff->prm.rcutl + ff->prm.skin + 0.3);
}

/* pressure coupling (md.rs:1116-1129, 1259-1277): PressureCalculation
 * (pressure_calculation.rs) on the atomic or molecular virial and
 * BerendsenBarostat (berendsen_barostat.rs), isotropic, rectangular or
 * triclinic boxes (every box vector scaled, the volume the determinant);
 * pg_last: the last atom of each pressure group (solute groups, then every
 * solvent molecule) */
int or_md_set_barostat(or_md *md, double p_ref, double comp, double tau_p, int vtype, int n_pg,
                       const int32_t *pg_last) {
    if ((md->ff->pbc.kind != 1 && md->ff->pbc.kind != 2) || vtype < 0 || vtype > 2 || n_pg < 0) return -1;
    md->baro = 1;
    md->p_ref = p_ref;
    md->comp = comp;
    md->tau_p = tau_p;
    md->vtype = vtype;
    md->n_pg = n_pg;
    free(md->pg_start);
    md->pg_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_pg + 1));
    md->pg_start[0] = 0;
    for (int g = 0; g < n_pg; ++g) md->pg_start[g + 1] = pg_last[g] + 1;
    md->ff->prm.virial = 1;
    return 0;
}

double or_md_pressure(const or_md *md) { return md->pressure; }
void or_md_box(const or_md *md, double *L) {
    for (int a = 0; a < 3; ++a) L[a] = md->ff->pbc.L[a];
}

/* prepare_virial's kinetic energy tensor (forcefield.rs:285-296, 1064-1120):
 * molecular, 0.5 (sum m v)(sum m v)^T / M per pressure group, or atomic,
 * 0.5 m v v^T per atom; ke[3 a + b] */
static void or_ke_tensor(const or_md *md, const double *vel, double *ke) {
    const double *mass = md->t->mass;
    for (int k = 0; k < 9; ++k) ke[k] = 0.0;
    if (md->vtype == 2 && md->n_pg > 0) {
        for (int g = 0; g < md->n_pg; ++g) {
            double mv[3] = {0.0, 0.0, 0.0}, m_tot = 0.0;
            for (int i = md->pg_start[g]; i < md->pg_start[g + 1]; ++i) {
                m_tot += mass[i];
                for (int a = 0; a < 3; ++a) mv[a] += mass[i] * vel[3 * i + a];
            }
            if (m_tot > 0.0)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) ke[3 * a + b] += 0.5 * mv[a] * mv[b] / m_tot;
        }
    } else {
        for (int k = 0; k < md->t->n_atoms; ++k)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) ke[3 * a + b] += 0.5 * mass[k] * vel[3 * k + a] * vel[3 * k + b];
    }
}

/* apply_molecular_virial_correction (forcefield.rs:1128-1200): per pressure
 * group the centre of mass of the chain-gathered positions (each atom's image
 * nearest the previous gathered one), then vir[b][a] -= f_i[a] r_i[b] with r_i
 * the image of atom i nearest that centre */
static void or_molecular_virial(const or_md *md, const double *pos, const double *force, double *vir) {
    const double *mass = md->t->mass;
    const or_pbc *pbc = &md->ff->pbc;
    double corr[9] = {0.0};
    for (int g = 0; g < md->n_pg; ++g) {
        const int i0 = md->pg_start[g], i1 = md->pg_start[g + 1];
        double com[3] = {0.0, 0.0, 0.0}, m_tot = 0.0, prev[3];
        for (int a = 0; a < 3; ++a) prev[a] = pos[3 * i0 + a];
        for (int i = i0; i < i1; ++i) {
            const double m = mass[i];
            double p[3], gathered[3];
            m_tot += m;
            or_nearest_image(pbc, pos + 3 * i, prev, p);
            for (int a = 0; a < 3; ++a) {
                gathered[a] = p[a] + prev[a];
                com[a] += m * gathered[a];
                prev[a] = gathered[a];
            }
        }
        if (m_tot > 0.0)
            for (int a = 0; a < 3; ++a) com[a] /= m_tot;
        for (int i = i0; i < i1; ++i) {
            double r[3];
            or_nearest_image(pbc, pos + 3 * i, com, r);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) corr[3 * b + a] += force[3 * i + a] * r[b];
        }
    }
    for (int k = 0; k < 9; ++k) vir[k] = vir[k] - corr[k];
}

/* shake (constraints.rs:213-389) with shake_one_constraint_full
 * (constraints.rs:501-570): sweeps over the constraints in order (the
 * solute's, then every solvent molecule's) until a sweep changes nothing, an
 * atom's constraints skipped in a sweep when no constraint moved it in the
 * previous one; each unconverged constraint (|l^2 - r^2| >= 2 tol l^2) moves
 * both atoms along the old bond vector by lambda = (l^2 - r^2) / (2 (old . r)
 * (1/m_i + 1/m_j)); then v = (x_new - x_old) / dt for every constrained atom.
 * vir (optional) += old bond (x) old bond * lambda / dt^2.  Returns the sweeps
 * done, or -1 when max_iter sweeps did not converge. */
int or_shake(int n_atoms, const double *inv_mass, int nc, const int32_t *ci, const int32_t *cj, const double *len,
             double tol, int max_iter, double dt, const double *pos_old, double *pos_new, double *vel_new,
             double *vir) {
    const double dt2 = dt * dt;
    unsigned char *skip_now = (unsigned char *)calloc((size_t)n_atoms, 1);
    unsigned char *skip_next = (unsigned char *)malloc((size_t)n_atoms);
    memset(skip_next, 1, (size_t)n_atoms);
    int converged = 0, it = 0;
    while (it < max_iter && !converged) {
        converged = 1;
        ++it;
        for (int k = 0; k < nc; ++k) {
            const int i = ci[k], j = cj[k];
            if (skip_now[i] && skip_now[j]) continue;
            if (inv_mass[i] == 0.0 && inv_mass[j] == 0.0) continue;
            const double l2 = len[k] * len[k];
            double r[3], rr[3];
            for (int a = 0; a < 3; ++a) {
                r[a] = pos_new[3 * i + a] - pos_new[3 * j + a];
                rr[a] = pos_old[3 * i + a] - pos_old[3 * j + a];
            }
            const double dist2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const double diff = l2 - dist2;
            if (fabs(diff) < l2 * tol * 2.0) continue;
            const double sp = rr[0] * r[0] + rr[1] * r[1] + rr[2] * r[2];
            const double ims = inv_mass[i] + inv_mass[j];
            if (ims < 1e-20) continue;
            const double lambda = diff / (sp * 2.0 * ims);
            if (vir) {
                const double lo = lambda / dt2;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) vir[3 * a + b] += rr[a] * rr[b] * lo;
            }
            for (int a = 0; a < 3; ++a) {
                const double corr = rr[a] * lambda;
                pos_new[3 * i + a] += corr * inv_mass[i];
                pos_new[3 * j + a] -= corr * inv_mass[j];
            }
            skip_next[i] = 0;
            skip_next[j] = 0;
            converged = 0;
        }
        unsigned char *t = skip_now;
        skip_now = skip_next;
        skip_next = t;
        memset(skip_next, 1, (size_t)n_atoms);
    }
    for (int k = 0; k < nc; ++k)
        for (int s = 0; s < 2; ++s) {
            const int a0 = s ? cj[k] : ci[k];
            for (int a = 0; a < 3; ++a) vel_new[3 * a0 + a] = (pos_new[3 * a0 + a] - pos_old[3 * a0 + a]) * (1.0 / dt);
        }
    free(skip_now);
    free(skip_next);
    return converged ? it : -1;
}

/* setup_lincs (constraints.rs:1168-1220) and lincs (constraints.rs:1241-1390),
 * LINCS of Hess et al. (1997) as gromosXX's lincs.cc writes it: constraint k
 * joins atoms ci[k] + off and cj[k] + off.
 *   setup: sdiag[k] = 1 / sqrt(1/m_i + 1/m_j); constraints a < b sharing an atom
 *   are coupled with coef = (1/m_common) sdiag[a] sdiag[b], negated when the
 *   shared atom is the first of both or the second of both; each constraint's
 *   coupled list ascends.
 *   B[k] = (x_old[j] - x_old[i]) / |...| (zero below 1e-10 nm); A[k][n] =
 *   coef[k][n] (B[k] . B[n]); rhs = sol = sdiag (B . (x[j] - x[i]) - l); order
 *   times: rhs' = A rhs, sol += rhs'; x[i] += B sdiag sol / m_i, x[j] -= B sdiag
 *   sol / m_j.  Then the rotational correction: p = sqrt(2 l^2 - |x[j] - x[i]|^2)
 *   (0 when not positive), rhs = sol = sdiag (l - p), the same expansion and
 *   update; v = (x - x_old) / dt for each constrained atom.  No virial.
 * Returns the largest relative length error after the solve. */
double or_lincs(int nc, const int32_t *ci, const int32_t *cj, const double *len, int off, int order,
                const double *inv_mass, double dt, const double *pos_old, double *pos_new, double *vel_new) {
    if (nc <= 0) return 0.0;
    double *sdiag = (double *)malloc(sizeof(double) * (size_t)nc);
    double *coef = (double *)calloc((size_t)nc * (size_t)nc, sizeof(double));   /* dense, symmetric */
    unsigned char *cpl = (unsigned char *)calloc((size_t)nc * (size_t)nc, 1);
    double *b = (double *)malloc(sizeof(double) * 3 * (size_t)nc);
    double *rhs0 = (double *)malloc(sizeof(double) * (size_t)nc);
    double *rhs1 = (double *)malloc(sizeof(double) * (size_t)nc);
    double *sol = (double *)malloc(sizeof(double) * (size_t)nc);
    double *rhs[2] = {rhs0, rhs1};
    for (int k = 0; k < nc; ++k) sdiag[k] = 1.0 / sqrt(inv_mass[ci[k] + off] + inv_mass[cj[k] + off]);
    for (int a = 0; a < nc; ++a)
        for (int c = a + 1; c < nc; ++c) {
            int common = -1;
            if (ci[a] == ci[c]) common = ci[a];
            else if (cj[a] == ci[c]) common = cj[a];
            else if (ci[a] == cj[c]) common = ci[a];
            else if (cj[a] == cj[c]) common = cj[a];
            if (common < 0) continue;
            double v = inv_mass[common + off] * sdiag[a] * sdiag[c];
            if (ci[a] == ci[c] || cj[a] == cj[c]) v *= -1.0;
            coef[(size_t)a * nc + c] = coef[(size_t)c * nc + a] = v;
            cpl[(size_t)a * nc + c] = cpl[(size_t)c * nc + a] = 1;
        }
    for (int k = 0; k < nc; ++k) {
        const int i = ci[k] + off, j = cj[k] + off;
        double r[3];
        for (int a = 0; a < 3; ++a) r[a] = pos_old[3 * j + a] - pos_old[3 * i + a];
        const double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int a = 0; a < 3; ++a) b[3 * k + a] = rl < 1e-10 ? 0.0 : r[a] / rl;
    }
#define OR_BDOT(p, q) (b[3 * (p)] * b[3 * (q)] + b[3 * (p) + 1] * b[3 * (q) + 1] + b[3 * (p) + 2] * b[3 * (q) + 2])
    for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < nc; ++k) {
            const int i = ci[k] + off, j = cj[k] + off;
            double r[3];
            for (int a = 0; a < 3; ++a) r[a] = pos_new[3 * j + a] - pos_new[3 * i + a];
            if (pass == 0) {
                const double proj = b[3 * k] * r[0] + b[3 * k + 1] * r[1] + b[3 * k + 2] * r[2];
                rhs[0][k] = sdiag[k] * (proj - len[k]);
            } else {
                const double diff = 2.0 * len[k] * len[k] - (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const double p = diff > 0.0 ? sqrt(diff) : 0.0;
                rhs[0][k] = sdiag[k] * (len[k] - p);
            }
            sol[k] = rhs[0][k];
        }
        int w = 1;
        for (int rec = 0; rec < order; ++rec) {
            for (int k = 0; k < nc; ++k) {
                rhs[w][k] = 0.0;
                for (int c = 0; c < nc; ++c)
                    if (cpl[(size_t)k * nc + c]) rhs[w][k] += coef[(size_t)k * nc + c] * OR_BDOT(k, c) * rhs[1 - w][c];
                sol[k] += rhs[w][k];
            }
            w = 1 - w;
        }
        for (int k = 0; k < nc; ++k) {
            const int i = ci[k] + off, j = cj[k] + off;
            const double corr = sdiag[k] * sol[k];
            const double si = corr * inv_mass[i], sj = corr * inv_mass[j];
            for (int a = 0; a < 3; ++a) {
                pos_new[3 * i + a] += b[3 * k + a] * si;
                pos_new[3 * j + a] -= b[3 * k + a] * sj;
            }
        }
    }
#undef OR_BDOT
    double err = 0.0;
    for (int k = 0; k < nc; ++k) {
        const int i = ci[k] + off, j = cj[k] + off;
        for (int a = 0; a < 3; ++a) {
            vel_new[3 * i + a] = (pos_new[3 * i + a] - pos_old[3 * i + a]) * (1.0 / dt);
            vel_new[3 * j + a] = (pos_new[3 * j + a] - pos_old[3 * j + a]) * (1.0 / dt);
        }
        double r[3];
        for (int a = 0; a < 3; ++a) r[a] = pos_new[3 * j + a] - pos_new[3 * i + a];
        const double e = fabs((sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) - len[k]) / len[k]);
        if (e > err) err = e;
    }
    free(sdiag);
    free(coef);
    free(cpl);
    free(b);
    free(rhs0);
    free(rhs1);
    free(sol);
    return err;
}

/* LINCS in the md step (md.rs:1244-1257, lincs_buffered constraints.rs:1397-1445),
 * after SHAKE and SETTLE: the solute's constraints (order_u > 0), then every
 * solvent molecule by the template (order_v > 0) */
int or_md_set_lincs(or_md *md, int order_u, int n_solute_c, const int32_t *sol_ij, const double *sol_len,
                    int order_v, int n_template, const int32_t *tmpl_ij, const double *tmpl_len) {
    if (n_template > OR_LINCS_TPL_MAX) return -1;
    free(md->lc_ij);
    free(md->lc_len);
    md->lc_ij = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(n_solute_c > 0 ? n_solute_c : 1));
    md->lc_len = (double *)malloc(sizeof(double) * (size_t)(n_solute_c > 0 ? n_solute_c : 1));
    for (int c = 0; c < n_solute_c; ++c) {
        md->lc_ij[2 * c] = sol_ij[2 * c];
        md->lc_ij[2 * c + 1] = sol_ij[2 * c + 1];
        md->lc_len[c] = sol_len[c];
    }
    md->lc_nu = order_u > 0 ? n_solute_c : 0;
    md->lc_nt = order_v > 0 ? n_template : 0;
    for (int c = 0; c < md->lc_nt; ++c) {
        md->lc_tij[2 * c] = tmpl_ij[2 * c];
        md->lc_tij[2 * c + 1] = tmpl_ij[2 * c + 1];
        md->lc_tlen[c] = tmpl_len[c];
    }
    md->lc_order_u = order_u;
    md->lc_order_v = order_v;
    return 0;
}

/* SHAKE in the md step (md.rs:1208-1236): the solute constraints (NTC 2: bonds
 * with a hydrogen, 3: all bonds; chosen by the caller) and, for NTCS shake,
 * every solvent molecule's template constraints; tolerance NTCP0 / NTCS0 */
int or_md_set_shake(or_md *md, int n_solute_c, const int32_t *sol_ij, const double *sol_len, int n_template,
                    const int32_t *tmpl_ij, const double *tmpl_len, double tol) {
    const or_topo *t = md->t;
    const int aps = t->atoms_per_solvent;
    const int nsm = aps > 0 ? (t->n_atoms - t->n_solute_atoms) / aps : 0;
    const int n = n_solute_c + (n_template > 0 ? nsm * n_template : 0);
    free(md->sh_i);
    free(md->sh_j);
    free(md->sh_len);
    md->sh_i = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    md->sh_j = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    md->sh_len = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int k = 0;
    for (int c = 0; c < n_solute_c; ++c, ++k) {
        md->sh_i[k] = sol_ij[2 * c];
        md->sh_j[k] = sol_ij[2 * c + 1];
        md->sh_len[k] = sol_len[c];
    }
    if (n_template > 0)
        for (int m = 0; m < nsm; ++m)
            for (int c = 0; c < n_template; ++c, ++k) {
                const int base = t->n_solute_atoms + m * aps;
                md->sh_i[k] = base + tmpl_ij[2 * c];
                md->sh_j[k] = base + tmpl_ij[2 * c + 1];
                md->sh_len[k] = tmpl_len[c];
            }
    md->n_shake = k;
    md->sh_tol = tol;
    return 0;
}

/* RemoveCOMMotion::new (remove_com_motion.rs:62-85): NTICOM >= 1 removes the
 * translation at step 0, >= 2 the rotation too; NSCM > 0 removes the
 * translation every NSCM steps, NSCM < 0 translation and rotation every |NSCM| */
void or_md_set_com(or_md *md, int nticom, int nscm) {
    md->nticom = nticom;
    md->nscm = nscm;
}

/* remove_com_translation (remove_com_motion.rs:90-113): v -= sum m v / sum m */
static void or_com_translation(int n, const double *mass, double *vel) {
    double m_tot = 0.0, cv[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        m_tot += mass[i];
        for (int a = 0; a < 3; ++a) cv[a] += mass[i] * vel[3 * i + a];
    }
    for (int a = 0; a < 3; ++a) cv[a] /= m_tot;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) vel[3 * i + a] -= cv[a];
}

/* remove_com_rotation (remove_com_motion.rs:129-240): positions of a rectangular
 * box put into (-L/2, L/2] (put_into_box_rect, remove_com_motion.rs:115-123),
 * taken back half a step (r - dt/2 v) about their centre of mass; the angular
 * velocity omega = I^-1 L (adjugate over the determinant, zero below 1e-12)
 * and v -= omega x r */
static void or_com_rotation(int n, const double *mass, const double *pos, double *vel, const or_pbc *pbc,
                            double dt) {
    const int wrap = pbc->kind == 1;
    double m_tot = 0.0, cv[3] = {0.0, 0.0, 0.0}, cr[3] = {0.0, 0.0, 0.0};
#define OR_WRAP(i, a) (wrap ? pos[3 * (i) + (a)] - round(pos[3 * (i) + (a)] / pbc->L[a]) * pbc->L[a] : pos[3 * (i) + (a)])
    for (int i = 0; i < n; ++i) {
        const double m = mass[i];
        m_tot += m;
        for (int a = 0; a < 3; ++a) {
            cv[a] += m * vel[3 * i + a];
            cr[a] += m * OR_WRAP(i, a) - 0.5 * m * vel[3 * i + a] * dt;
        }
    }
    for (int a = 0; a < 3; ++a) {
        cv[a] /= m_tot;
        cr[a] /= m_tot;
    }
    double L[3] = {0.0, 0.0, 0.0};
    double i00 = 0.0, i11 = 0.0, i22 = 0.0, i01 = 0.0, i02 = 0.0, i12 = 0.0;
    for (int i = 0; i < n; ++i) {
        const double m = mass[i];
        double r[3], u[3];
        for (int a = 0; a < 3; ++a) {
            r[a] = OR_WRAP(i, a) - 0.5 * dt * vel[3 * i + a] - cr[a];
            u[a] = vel[3 * i + a] - cv[a];
        }
        L[0] += m * (r[1] * u[2] - r[2] * u[1]);
        L[1] += m * (r[2] * u[0] - r[0] * u[2]);
        L[2] += m * (r[0] * u[1] - r[1] * u[0]);
        i00 += m * (r[1] * r[1] + r[2] * r[2]);
        i11 += m * (r[0] * r[0] + r[2] * r[2]);
        i22 += m * (r[0] * r[0] + r[1] * r[1]);
        i01 -= m * r[0] * r[1];
        i02 -= m * r[0] * r[2];
        i12 -= m * r[1] * r[2];
    }
    const double den = -i02 * i02 * i11 + 2.0 * i01 * i02 * i12 - i00 * i12 * i12 - i01 * i01 * i22 + i00 * i11 * i22;
    const double a00 = -i12 * i12 + i11 * i22, a01 = i02 * i12 - i01 * i22, a02 = -i02 * i11 + i01 * i12;
    const double a11 = -i02 * i02 + i00 * i22, a12 = i01 * i02 - i00 * i12, a22 = -i01 * i01 + i00 * i11;
    double w[3] = {0.0, 0.0, 0.0};
    if (!(den < 1e-12)) {
        w[0] = (a00 * L[0] + a01 * L[1] + a02 * L[2]) / den;
        w[1] = (a01 * L[0] + a11 * L[1] + a12 * L[2]) / den;
        w[2] = (a02 * L[0] + a12 * L[1] + a22 * L[2]) / den;
    }
    for (int i = 0; i < n; ++i) {
        double r[3];
        for (int a = 0; a < 3; ++a) r[a] = OR_WRAP(i, a) - 0.5 * dt * vel[3 * i + a] - cr[a];
        vel[3 * i] -= w[1] * r[2] - w[2] * r[1];
        vel[3 * i + 1] -= w[2] * r[0] - w[0] * r[2];
        vel[3 * i + 2] -= w[0] * r[1] - w[1] * r[0];
    }
#undef OR_WRAP
}

/* RemoveCOMMotion::apply (remove_com_motion.rs:246-313) at md step `step` */
void or_remove_com(int n, const double *mass, const double *pos, double *vel, const or_pbc *pbc, double dt,
                   int nticom, int nscm, int64_t step) {
    int trans, rot;
    const int64_t skip = nscm > 0 ? nscm : -(int64_t)nscm;
    if (step == 0) {
        trans = nticom >= 1;
        rot = nticom >= 2;
    } else if (skip > 0 && step % skip == 0) {
        trans = 1;
        rot = nscm < 0 && pbc->kind == 0;   /* periodic boxes: no periodic rotation removal */
    } else {
        return;
    }
    if (trans) or_com_translation(n, mass, vel);
    if (rot) or_com_rotation(n, mass, pos, vel, pbc, dt);
}

/* BerendsenThermostat::apply (berendsen_thermostat.rs:61-100): the scale from
 * E_kin_new = 0.5 sum m |v|^2 of the velocities entering the step (stored by the
 * previous TemperatureCalculation, temperature.rs:76-85, or by its init) */
double or_berendsen_scale(int n, const double *mass, const double *vel, double t_ref, double tau, double dof,
                          double dt) {
    double ekin = 0.0;
    for (int i = 0; i < n; ++i)
        ekin += 0.5 * mass[i] * (vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] + vel[3 * i + 2] * vel[3 * i + 2]);
    return or_berendsen_scale_ekin(ekin, t_ref, tau, dof, dt);
}

double or_berendsen_scale_ekin(double ekin, double t_ref, double tau, double dof, double dt) {
    if (tau < 0.0) return 1.0;
    double free_temp = dof > 0.0 ? 2.0 * ekin / (dof * OR_KB) : 0.0;
    if (free_temp < 1e-10) free_temp = t_ref;
    if (free_temp < 1e-10) return 1.0;
    if (tau == 0.0) return sqrt(t_ref / free_temp);
    return sqrt(1.0 + dt / tau * (t_ref / free_temp - 1.0));
}

int or_md_run(or_md *md, int nsteps, double *energies) {
    int n = md->t->n_atoms;
    for (int s = 0; s < nsteps; ++s) {
        int c = md->cur, o = 1 - c;
        double elj, ecrf, dh;
        /* the bath reads E_kin_new stored by the previous step, before this
         * step's centre-of-mass removal (md.rs:1095-1097 runs that first) */
        double ek_bath = 0.0;
        for (int i = 0; i < n; ++i) {
            const double *v = md->vel[c] + 3 * i;
            ek_bath += 0.5 * md->t->mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        if (md->nticom >= 1 || md->nscm != 0)
            or_remove_com(n, md->t->mass, md->pos[c], md->vel[c], &md->ff->pbc, md->dt, md->nticom, md->nscm,
                          md->step);
        /* prepare_virial: the kinetic energy tensor of the velocities entering
         * the step, before the forces (forcefield.rs:285-296) */
        double ke[9], vir[9];
        if (md->baro) or_ke_tensor(md, md->vel[c], ke);
        or_ff_apply(md->ff, md->pos[c], md->force[c], &elj, &ecrf, md->baro ? vir : NULL, &dh);
        if (md->baro && md->vtype == 2 && md->n_pg > 0) or_molecular_virial(md, md->pos[c], md->force[c], vir);
        md->pairs += or_ff_last_pair_count(md->ff);
        /* exchange_state: the current buffer becomes old */
        if (md->thermo) {
            /* LeapFrogVelocity, thermostat scaling, LeapFrogPosition (md.rs:1143-1206) */
            double lam;
            if (md->thermo == 1)
                lam = or_berendsen_scale_ekin(ek_bath, md->t_ref, md->tau_t, md->dof_t, md->dt);
            else
                lam = or_nosehoover_step(md->zeta, md->nhc, ek_bath, md->t_ref, md->tau_t, md->dof_t, md->dt);
            for (int i = 0; i < n; ++i)
                for (int a = 0; a < 3; ++a) {
                    double acc = md->force[c][3 * i + a] * md->inv_mass[i];
                    md->vel[o][3 * i + a] = (md->vel[c][3 * i + a] + acc * md->dt) * lam;
                }
            for (int i = 0; i < n; ++i)
                for (int a = 0; a < 3; ++a) md->pos[o][3 * i + a] = md->pos[c][3 * i + a] + md->vel[o][3 * i + a] * md->dt;
        } else {
            or_leapfrog(n, md->inv_mass, md->dt, md->pos[c], md->vel[c], md->force[c], md->pos[o], md->vel[o]);
        }
        if (md->n_shake > 0) {
            int it = or_shake(n, md->inv_mass, md->n_shake, md->sh_i, md->sh_j, md->sh_len, md->sh_tol, 1000, md->dt,
                              md->pos[c], md->pos[o], md->vel[o], md->baro ? vir : NULL);
            if (it < 0) return -1;
        }
        if (md->use_settle) {
            /* the constraint's atomic virial joins the pressure's virial after the
             * molecular correction (constraints.rs:978-1010) */
            int err = or_settle(md->t->n_solute_atoms, n, md->m_o, md->m_h, md->d_oh, md->d_hh, md->dt,
                                md->pos[c], md->pos[o], md->vel[o], NULL, md->baro ? vir : NULL);
            if (err) return -1;
        }
        if (md->lc_nu > 0 || md->lc_nt > 0) {
            /* the solute's constraints, then each solvent molecule's (split
             * index pairs for or_lincs) */
            int32_t *ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(md->lc_nu + md->lc_nt + 1));
            int32_t *cj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(md->lc_nu + md->lc_nt + 1));
            for (int k = 0; k < md->lc_nu; ++k) {
                ci[k] = md->lc_ij[2 * k];
                cj[k] = md->lc_ij[2 * k + 1];
            }
            or_lincs(md->lc_nu, ci, cj, md->lc_len, 0, md->lc_order_u, md->inv_mass, md->dt, md->pos[c], md->pos[o],
                     md->vel[o]);
            const int aps = md->t->atoms_per_solvent;
            const int nsm = aps > 0 ? (n - md->t->n_solute_atoms) / aps : 0;
            for (int k = 0; k < md->lc_nt; ++k) {
                ci[k] = md->lc_tij[2 * k];
                cj[k] = md->lc_tij[2 * k + 1];
            }
            if (md->lc_nt > 0)
                for (int m = 0; m < nsm; ++m)
                    or_lincs(md->lc_nt, ci, cj, md->lc_tlen, md->t->n_solute_atoms + m * aps, md->lc_order_v,
                             md->inv_mass, md->dt, md->pos[c], md->pos[o], md->vel[o]);
            free(ci);
            free(cj);
        }
        double ek = or_ekin(n, md->t->mass, md->vel[o], md->vel[c]);
        if (md->baro) {
            /* PressureCalculation (pressure_calculation.rs:40-80) on this step's
             * tensors and box: P = (KE + vir / 2) * 2 / V; BerendsenBarostat
             * (berendsen_barostat.rs:30-70): mu = (1 - comp dt / tau (P0 - P))^(1/3)
             * scales the box and the new positions */
            const double *L = md->ff->pbc.L;
            const or_pbc *pb = &md->ff->pbc;
            /* BoxConfig::volume, the box matrix's determinant c . (a x b) (glam's
             * order, configuration.rs:86-89); the rectangular product equals it */
            const double V = pb->kind == 2
                                 ? pb->c[0] * (pb->a[1] * pb->b[2] - pb->b[1] * pb->a[2]) +
                                       pb->c[1] * (pb->a[2] * pb->b[0] - pb->b[2] * pb->a[0]) +
                                       pb->c[2] * (pb->a[0] * pb->b[1] - pb->b[0] * pb->a[1])
                                 : L[0] * L[1] * L[2];
            if (md->vtype != 0 && !(V < 1e-30)) {
                const double factor = 2.0 / V;
                double pt[3];
                for (int a = 0; a < 3; ++a) pt[a] = (ke[4 * a] + vir[4 * a] * 0.5) * factor;
                md->pressure = (pt[0] + pt[1] + pt[2]) / 3.0;
            }
            const double mu = pow(1.0 - md->comp * md->dt / md->tau_p * (md->p_ref - md->pressure), 1.0 / 3.0);
            if (pb->kind == 2) {
                /* every box vector scaled (berendsen_barostat.rs:40-48) */
                double v9[9];
                for (int a = 0; a < 3; ++a) {
                    v9[a] = pb->a[a] * mu;
                    v9[3 + a] = pb->b[a] * mu;
                    v9[6 + a] = pb->c[a] * mu;
                }
                or_pbc_init(&md->ff->pbc, 2, v9);   /* CG-grouped loops stay off for triclinic */
            } else {
                double Ln[3] = {L[0] * mu, L[1] * mu, L[2] * mu};
                or_ff_set_box(md->ff, Ln);
            }
            for (int k = 0; k < 3 * n; ++k) md->pos[o][k] *= mu;
        }
        if (energies) {
            energies[5 * s + 0] = elj;
            energies[5 * s + 1] = ecrf;
            /* update_potential_total (configuration.rs:197-207): bond_total holds every
             * covalent term; the other covalent totals stay zero */
            const or_bonded_e *be = &md->ff->be;
            double eb = ((be->e_bond + be->e_angle) + be->e_dihedral) + be->e_improper;
            energies[5 * s + 2] = (eb + elj) + ecrf;
            energies[5 * s + 3] = ek;
            energies[5 * s + 4] = dh;
        }
        md->cur = o;
        md->step++;
    }
    return 0;
}

void or_md_get(const or_md *md, double *pos, double *vel, double *force) {
    size_t b = sizeof(double) * 3 * (size_t)md->t->n_atoms;
    if (pos) memcpy(pos, md->pos[md->cur], b);
    if (vel) memcpy(vel, md->vel[md->cur], b);
    if (force) memcpy(force, md->force[1 - md->cur], b);
}

int64_t or_md_pair_count(const or_md *md) { return md->pairs; }

static int g_threads = 0;
int or_num_threads(void) {
    if (g_threads > 0) return g_threads;
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int n) { g_threads = n; }
