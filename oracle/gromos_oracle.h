/*
 * CPU oracle: a plain-C restatement of the gromos-rs nonbonded hot path.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the
 * smoke check and the CPU-baseline leg of bench.py may load it, and only as
 * the checker (or the timed CPU baseline).  The GPU library never links it.
 *
 * Every function states the reference file:line it follows (paths relative to
 * the gromos-rs repository root).  Arithmetic is FP64, written in the same
 * operation order as the Rust source, and compiled with -ffp-contract=off so
 * that no multiply-add is fused (Rust never fuses).
 *
 * Parity status (how each part is pinned; see tests/test_oracle_*.py):
 *   - pair kernel, CRF constants, PBC, CG pairlist, RF excluded/self, 1-4,
 *     solvent loop, twin range, leap-frog, SETTLE, E_kin: pinned by the
 *     gromosXX double-precision fixtures the reference ships, by closed forms
 *     and by brute-force / invariant checks.
 *   - soft-core LJ (zero charge): pinned by the ch4_water_fep fixture.
 *   - soft-core CRF with charges, dH/dlambda, perturbed self / excluded / 1-4 /
 *     atom-pair corrections: PARITY UNPINNED (no fixture or runnable
 *     reference covers them); checked only by finite differences and limits.
 */
#ifndef GROMOS_ORACLE_H
#define GROMOS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- boundary conditions (crates/gromos-core/src/math.rs:20-74) ---- */
typedef struct {
    int kind;          /* 0 = vacuum, 1 = rectangular, 2 = triclinic */
    double L[3];       /* box edge lengths (triclinic: the diagonal ax, by, cz) */
    double invL[3];    /* 1.0 / L, computed once (math.rs:49-55) */
    double half[3];
    double a[3], b[3], c[3];   /* triclinic box vectors, lower triangular (math.rs:75-92) */
} or_pbc;

/* box: 3 edge lengths, or for kind 2 the vectors a, b, c (9 values) */
void or_pbc_init(or_pbc *p, int kind, const double *box);
void or_nearest_image(const or_pbc *p, const double *ri, const double *rj, double *r);

/* ---- CRF constants (crates/gromos-forces/src/nonbonded/params.rs:60-97) ---- */
typedef struct {
    double crf_cut;
    double crf_2cut3i;
    double crf_cut3i;
    double cutoff_sq;
} or_crf;

void or_crf_new(double cutoff, double epsilon, double rf_epsilon, double rf_kappa, or_crf *out);

/* ---- pair kernel (innerloops.rs:22-55) ---- */
void or_lj_crf_interaction(const double r[3], double c6, double c12, double q_prod,
                           const or_crf *crf, double *f, double *e_lj, double *e_crf);

/* ---- covalent terms of the solute (crates/gromos-forces/src/bonded) ---- */
typedef struct {
    int32_t nb;  const int32_t *bonds;     const double *bond_types;      /* i j t; kq kh r0 */
    int32_t na;  const int32_t *angles;    const double *angle_types;     /* i j k t; kc kh theta0 */
    int32_t ni;  const int32_t *impropers; const double *improper_types;  /* i j k l t; k q0 */
    int32_t nd;  const int32_t *dihedrals; const double *dihedral_types;  /* i j k l t; k cospd m */
    int32_t nbt, nat, nit, ndt;            /* type table sizes (terms of a larger type are skipped) */
    int32_t ntf[4];                        /* FORCE block NTF(1..4): bond, angle, improper, dihedral on */
} or_bonded;

/* e: bond, angle, dihedral, improper energy */
typedef struct { double e_bond, e_angle, e_dihedral, e_improper; } or_bonded_e;

/* calculate_bonded_forces_ntf with quartic bonds (bonded/mod.rs:141-185): each
 * kind into its own zeroed buffer, summed per atom in the order bonds, angles,
 * dihedrals, impropers; f (3n) and vir (9, row a col b) are ASSIGNED */
void or_bonded_forces(const or_bonded *b, int n_atoms, const double *pos, double *f, double vir[9],
                      or_bonded_e *e);

/* ---- topology essentials ---- */
typedef struct {
    int32_t n_atoms;
    int32_t n_solute_atoms;
    int32_t n_cg;
    int32_t n_solute_cg;
    const int32_t *cg_start;     /* n_cg + 1 offsets; charge groups are contiguous */
    const double *charge;        /* n_atoms */
    const int32_t *iac;          /* n_atoms, 0-based type */
    const double *mass;          /* n_atoms */
    int32_t n_types;
    const double *lj;            /* n_types*n_types*4: c6, c12, cs6, cs12 (params.rs:6-13) */
    const int32_t *excl_off;     /* n_atoms + 1; sorted, symmetric (topology.rs 1-2/1-3/solvent) */
    const int32_t *excl;
    const int32_t *o14_off;      /* n_atoms + 1; per-atom 1-4 partners as read (INE14) */
    const int32_t *o14;
    int32_t atoms_per_solvent;
    const or_bonded *bonded;     /* NULL: no covalent terms */
} or_topo;

/* ---- growable u32 pair list ---- */
typedef struct {
    int64_t n;
    int64_t cap;
    uint32_t *p;   /* 2*n entries: (i, j) */
} or_plist;

typedef struct {
    or_plist solute_short, solute_long, solvent_short, solvent_long;
} or_pairlist;

void or_pairlist_free(or_pairlist *pl);
void or_pairlist_clear(or_pairlist *pl);

/* is_excluded_or_14 (crates/gromos-core/src/topology.rs:853-865) */
int or_is_excluded_or_14(const or_topo *t, int i, int j);

/* StandardPairlistAlgorithm::update_chargegroup_based (pairlist.rs:183-442) */
void or_pairlist_standard(const or_topo *t, const double *pos, const or_pbc *pbc,
                          double rcutp, double rcutl, double skin, or_pairlist *out);
/* CellListPairlistAlgorithm::update_cell_list (pairlist.rs:537-655, 657-777) */
void or_pairlist_cell(const or_topo *t, const double *pos, const or_pbc *pbc,
                      double rcutp, double rcutl, double skin, or_pairlist *out);

/* ---- force storage (params.rs:99-138) ---- */
typedef struct {
    int32_t n;
    double *f;        /* 3n */
    double e_lj, e_crf;
    double vir[9];    /* row a, col b: sum r[a]*f[b] */
} or_storage;

void or_storage_init(or_storage *s, int n);
void or_storage_free(or_storage *s);
void or_storage_clear(or_storage *s);
void or_storage_merge(or_storage *dst, const or_storage *src);

/* ---- inner loops (innerloops.rs:105-309) ---- */
void or_lj_crf_innerloop(const double *pos, const double *q, const int32_t *iac,
                         const uint32_t *pairs, int64_t np, const double *lj, int32_t n_types,
                         const or_crf *crf, const or_pbc *pbc, double fpepsi, int virial,
                         or_storage *s);
/* groups: 4 u32 per group: ref_atom_i, ref_atom_j, start, end (params.rs:140-154) */
void or_lj_crf_innerloop_cg_grouped(const double *pos, const double *q, const int32_t *iac,
                                    const uint32_t *pairs, const uint32_t *groups, int64_t ng,
                                    const double *lj, int32_t n_types, const or_crf *crf,
                                    const or_pbc *pbc, double fpepsi, int virial, or_storage *s);
void or_solvent_innerloop(const double *pos, const double *q, const int32_t *iac,
                          const uint32_t *pairs, int64_t np, const double *lj, int32_t n_types,
                          const or_crf *crf, const or_pbc *pbc, int32_t atoms_per_solvent,
                          double fpepsi, int virial, or_storage *s);

/* OpenMP fold/reduce variants mirroring the rayon loops (innerloops.rs:430-864):
 * per-thread storage, the same serial thresholds and chunk sizes. */
void or_lj_crf_innerloop_parallel(const double *pos, const double *q, const int32_t *iac,
                                  const uint32_t *pairs, int64_t np, const double *lj,
                                  int32_t n_types, const or_crf *crf, const or_pbc *pbc,
                                  double fpepsi, int virial, or_storage *out);
void or_solvent_innerloop_parallel(const double *pos, const double *q, const int32_t *iac,
                                   const uint32_t *pairs, int64_t np, const double *lj,
                                   int32_t n_types, const or_crf *crf, const or_pbc *pbc,
                                   int32_t atoms_per_solvent, double fpepsi, int virial,
                                   or_storage *out);

/* ---- excluded / 1-4 (excluded.rs:25-206) ---- */
void or_rf_excluded_interactions(const or_topo *t, const double *pos, const or_crf *crf,
                                 const or_pbc *pbc, double fpepsi, or_storage *s);
void or_one_four_interaction_loop(const or_topo *t, const double *pos, const or_crf *crf,
                                  const or_pbc *pbc, double fpepsi, double coulomb_scaling,
                                  or_storage *s);

/* ---- perturbation (perturbed.rs) ---- */
typedef struct {
    double a_lj_lambda_n, b_lj_lambda_n, a_crf_lambda_n, b_crf_lambda_n;
    double a_lj_lambda_n_1, b_lj_lambda_n_1, a_crf_lambda_n_1, b_crf_lambda_n_1;
    double a_ljs_lambda, b_ljs_lambda, a_ljs_lambda2, b_ljs_lambda2;
    double a_crfs_lambda, b_crfs_lambda, a_crfs_lambda2, b_crfs_lambda2;
    int32_t lambda_exp;
} or_lambda;

void or_lambda_from(double lj_l, double ljs_l, double crf_l, double crfs_l, double lj_d,
                    double ljs_d, double crf_d, double crfs_d, int n, or_lambda *out);

/* perturbed_lj_crf_interaction (perturbed.rs:124-266): out[5] = f, e_lj, e_crf, de_lj, de_crf */
void or_perturbed_lj_crf_interaction(const double r[3], double a_c6, double a_c12, double b_c6,
                                     double b_c12, double a_q, double b_q, double alpha_lj,
                                     double alpha_crf, const or_lambda *lp, const or_crf *crf,
                                     double fpepsi, double out[5]);

/* per-atom perturbation info (perturbed.rs:271-280); flag 0 = not perturbed */
typedef struct {
    const uint8_t *flag;
    const int32_t *a_iac, *b_iac;
    const double *a_q, *b_q, *alpha_lj, *alpha_crf;
    /* PERTATOMPAIR list: i, j, a_type, b_type (b_type < 0 = absent in state B) */
    int32_t n_atom_pairs;
    const int32_t *atom_pairs;
} or_pert;

typedef struct {
    double *f;  /* 3n */
    double de_lj, de_crf, dhdl;
} or_pertcorr;

void or_perturbed_pairlist_correction(const double *pos, const uint32_t *pairs, int64_t np,
                                      const or_pert *pt, const double *q, const int32_t *iac,
                                      const double *lj, int32_t n_types, const or_crf *crf,
                                      const or_lambda *lp, const or_pbc *pbc, double fpepsi,
                                      int n_atoms, or_pertcorr *out);
void or_perturbed_self_energy_correction(const or_topo *t, const or_pert *pt, const or_crf *crf,
                                         const or_lambda *lp, double fpepsi, double *de,
                                         double *dhdl);
void or_perturbed_excluded_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                      const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                      double fpepsi, or_pertcorr *out);
void or_perturbed_one_four_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                      const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                      double fpepsi, or_pertcorr *out);
void or_perturbed_atom_pair_correction(const or_topo *t, const double *pos, const or_pert *pt,
                                       const or_crf *crf, const or_lambda *lp, const or_pbc *pbc,
                                       double fpepsi, or_pertcorr *out);

/* ---- Forcefield::apply (crates/gromos-integrators/src/algorithms/forcefield.rs:226-1060),
 *      nonbonded part only (bonded terms are outside this path) ---- */
typedef struct or_ff or_ff;

typedef struct {
    double rcutp, rcutl, skin;
    int32_t nsnb;            /* pairlist update frequency */
    int32_t algorithm;       /* 0 standard, 2 cell list (pairlist.rs:815-822) */
    double epsrf, kappa;     /* CRF built from rcutl, epsilon 1 (md.rs:326-329) */
    double fpepsi;
    int32_t parallel;        /* 1: OpenMP fold/reduce loops (md.rs:1113: N > 100) */
    int32_t virial;          /* 1: virial on the short-range loops */
    /* perturbation: enabled when pert.flag != NULL and some atom is perturbed */
    double lambda;
    int32_t nlam;
} or_ff_params;

or_ff *or_ff_create(const or_topo *t, const or_pbc *pbc, const or_ff_params *prm,
                    const or_pert *pt);
void or_ff_destroy(or_ff *ff);
/* Initial pairlist build before the step loop (md.rs:866-868). */
void or_ff_initial_pairlist(or_ff *ff, const double *pos);
/* One Forcefield::apply; writes forces (3n), energies, virial, dH/dlambda. */
int or_ff_apply(or_ff *ff, const double *pos, double *forces, double *e_lj, double *e_crf,
                double *vir, double *dhdl);
const or_pairlist *or_ff_pairlist(const or_ff *ff);
/* the covalent energies of the last or_ff_apply (zero without covalent terms) */
void or_ff_last_bonded(const or_ff *ff, or_bonded_e *e);
int or_ff_used_cg_grouped(const or_ff *ff);
/* number of pair interactions evaluated by the last apply (for throughput accounting) */
int64_t or_ff_last_pair_count(const or_ff *ff);

/* ---- step pieces ---- */
/* LeapFrogVelocity + LeapFrogPosition (leap_frog.rs:37-148) */
void or_leapfrog(int n, const double *inv_mass, double dt, const double *pos_old,
                 const double *vel_old, const double *f_old, double *pos_new, double *vel_new);
/* settle (crates/gromos-integrators/src/constraints.rs:815-1036); returns #failures.
 * pos_new and vel_new are corrected in place; cons_force (3n, optional) and vir (9,
 * optional) receive the constraint force and its virial. */
int or_settle(int n_solute, int n_atoms, double m_o, double m_h, double dist_oh, double dist_hh,
              double dt, const double *pos_old, double *pos_new, double *vel_new,
              double *cons_force, double *vir);
/* TemperatureCalculation (temperature.rs:63-87) */
double or_ekin(int n, const double *mass, const double *vel_new, const double *vel_old);

/* ---- full MD driver used for parity runs and the CPU baseline ---- */
typedef struct or_md or_md;
or_md *or_md_create(const or_topo *t, const or_pbc *pbc, const or_ff_params *prm,
                    const or_pert *pt, double dt, int use_settle, const double *pos,
                    const double *vel);
void or_md_destroy(or_md *md);
/* runs nsteps of: Forcefield, LeapFrogVelocity, LeapFrogPosition, [SETTLE],
 * TemperatureCalculation, EnergyCalculation (md.rs:1143-1286, 1545-1557).
 * energies: per step 5 doubles (e_lj, e_crf, e_pot, e_kin, dhdl), may be NULL. */
void or_md_set_settle_geometry(or_md *md, double d_oh, double d_hh);
int or_md_run(or_md *md, int nsteps, double *energies);
/* Berendsen weak coupling (berendsen_thermostat.rs), one bath over all atoms */
#define OR_KB 0.00831441   /* kJ/(mol K), units.rs:47 */
void or_md_set_berendsen(or_md *md, double t_ref, double tau, double dof);
/* Nose-Hoover chain (nosehoover_thermostat.rs), one bath over all atoms, chain
 * length 1..OR_NHC_MAX; returns -1 for another length */
#define OR_NHC_MAX 16
int or_md_set_nosehoover(or_md *md, double t_ref, double tau, double dof, int nhc);
double or_nosehoover_step(double *zeta, int nhc, double ekin, double t_ref, double tau, double dof, double dt);
void or_md_get_nosehoover(const or_md *md, double *zeta);
double or_berendsen_scale_ekin(double ekin, double t_ref, double tau, double dof, double dt);
/* Berendsen barostat with the pressure from the atomic (vtype 1) or molecular
 * (vtype 2) virial, rectangular boxes; pg_last: last atom of each pressure
 * group; returns -1 for another box or virial kind */
int or_md_set_barostat(or_md *md, double p_ref, double comp, double tau_p, int vtype, int n_pg,
                       const int32_t *pg_last);
double or_md_pressure(const or_md *md);
/* SHAKE (constraints.rs:213-570) in the md step: solute constraints (pairs,
 * lengths) then each solvent molecule's template constraints */
int or_md_set_shake(or_md *md, int n_solute_c, const int32_t *sol_ij, const double *sol_len, int n_template,
                    const int32_t *tmpl_ij, const double *tmpl_len, double tol);
int or_shake(int n_atoms, const double *inv_mass, int nc, const int32_t *ci, const int32_t *cj, const double *len,
             double tol, int max_iter, double dt, const double *pos_old, double *pos_new, double *vel_new,
             double *vir);
/* LINCS (constraints.rs:1168-1390) of constraints (ci[k] + off, cj[k] + off);
 * returns the largest relative length error after it */
#define OR_LINCS_TPL_MAX 6
double or_lincs(int nc, const int32_t *ci, const int32_t *cj, const double *len, int off, int order,
                const double *inv_mass, double dt, const double *pos_old, double *pos_new, double *vel_new);
/* LINCS in the md step after SHAKE / SETTLE: the solute's constraints with
 * expansion order order_u (0: none), then each solvent molecule's template
 * (at most OR_LINCS_TPL_MAX constraints) with order_v (0: none) */
int or_md_set_lincs(or_md *md, int order_u, int n_solute_c, const int32_t *sol_ij, const double *sol_len,
                    int order_v, int n_template, const int32_t *tmpl_ij, const double *tmpl_len);
void or_md_box(const or_md *md, double *L);
/* centre-of-mass motion removal (remove_com_motion.rs) at the start of each md step */
void or_md_set_com(or_md *md, int nticom, int nscm);
void or_remove_com(int n, const double *mass, const double *pos, double *vel, const or_pbc *pbc, double dt,
                   int nticom, int nscm, int64_t step);
double or_berendsen_scale(int n, const double *mass, const double *vel, double t_ref, double tau, double dof,
                          double dt);
void or_md_get(const or_md *md, double *pos, double *vel, double *force);
int64_t or_md_pair_count(const or_md *md);

int or_num_threads(void);
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
