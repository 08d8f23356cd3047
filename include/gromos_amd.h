/*
 * gromos_amd: MI355X (gfx950) nonbonded hot path for gromos-rs.
 *
 * C ABI: plain pointers and sizes.  Two levels:
 *
 *  1. Pair-list kernels.  Drop-ins for the gromos-forces free functions
 *     (crates/gromos-forces/src/nonbonded/{innerloops,excluded,perturbed}.rs).
 *     Every pointer is DEVICE memory; results ACCUMULATE into the caller's
 *     force / energy / virial buffers like the reference's `&mut ForceStorage`
 *     (callers clear first, forcefield.rs:418).  `stream` is a hipStream_t
 *     (NULL = default stream); calls are asynchronous on it.
 *
 *  2. Device-resident systems.  One or more MD jobs kept in HBM (positions,
 *     velocities, pair lists); the GPU builds the charge-group pair list
 *     (pairlist.rs:183-442 semantics) and runs Forcefield::apply
 *     (forcefield.rs:226-1060, nonbonded part), leap-frog, SETTLE and E_kin.
 *     Host pointers here are HOST memory; the library copies.
 *
 * Conventions (both levels):
 *   - return 0 on success, a negative gromos_amd_status on error; the message
 *     of the last error on the calling thread is gromos_amd_last_error();
 *   - positions / forces are (x, y, z) per atom, atom-major, f64;
 *   - the virial is row-major: vir[3*a + b] = sum r[a] * f[b]
 *     (innerloops.rs:199-207);
 *   - the LJ table is flat [ti * n_types + tj] x (c6, c12, cs6, cs12)
 *     (LJParamMatrix / LJParameters, params.rs:6-58); type indices are 0-based;
 *   - boundary: 0 = vacuum, 1 = rectangular (math.rs:20-74), 2 = triclinic
 *     (math.rs:75-141); a `box` argument holds 3 edge lengths, or for
 *     boundary 2 the box vectors a, b, c (9 doubles, lower triangular).
 * A system handle must be used from one host thread at a time.
 */
#ifndef GROMOS_AMD_H
#define GROMOS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gromos_amd_status {
    GROMOS_AMD_OK = 0,
    GROMOS_AMD_EINVAL = -1,     /* bad argument / shape */
    GROMOS_AMD_EHIP = -2,       /* HIP runtime error */
    GROMOS_AMD_ENOMEM = -3,     /* device allocation failed / over budget */
    GROMOS_AMD_ESETTLE = -4,    /* SETTLE failed (constraints.rs:896-951) */
    GROMOS_AMD_EBUSY = -5,      /* no admission (server full) */
    GROMOS_AMD_ENOTFOUND = -6,
    GROMOS_AMD_EFORCE = -7,     /* a non-finite or out-of-range force (|F| >= 2^25 kJ/mol/nm): the job diverged */
    GROMOS_AMD_EBOX = -8        /* a barostat shrank the box below the pair search's cell grid (5 % margin) */
};

/* CRFParameters (params.rs:60-68). */
typedef struct {
    double crf_cut;
    double crf_2cut3i;
    double crf_cut3i;
    double cutoff_sq;
} gromos_amd_crf;

const char *gromos_amd_last_error(void);
const char *gromos_amd_version(void);
/* number of visible HIP devices (0 when none) */
int gromos_amd_device_count(void);

/* CRFParameters::new (params.rs:78-97); pure host arithmetic. */
int gromos_amd_crf_new(double cutoff, double epsilon, double rf_epsilon, double rf_kappa,
                       gromos_amd_crf *out);

/* ------------------------------------------------------------------ */
/* 1. pair-list kernels (device pointers, accumulate semantics)         */
/* ------------------------------------------------------------------ */

/* lj_crf_innerloop[_novirial] (innerloops.rs:318-368): per-pair nearest image.
 * pairs: 2*n_pairs u32 (i, j).  energies: 2 doubles (e_lj, e_crf).
 * virial may be NULL when do_virial == 0.  Energies and virial are summed in a
 * fixed order through per-block partial sums (a scratch buffer from the
 * device's stream-ordered pool, hipMallocAsync on `stream`), so they are the
 * same bits on every call. */
int gromos_amd_lj_crf_innerloop(const double *pos, const double *charges, const uint32_t *iac,
                                const uint32_t *pairs, int64_t n_pairs, const double *lj,
                                int32_t n_types, const gromos_amd_crf *crf, int32_t boundary,
                                const double box[3], double four_pi_eps_i, int32_t do_virial,
                                double *forces, double *energies, double *virial, void *stream);

/* gromos_amd_lj_crf_innerloop without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-block partial sums live in a static device array of the library
 * (65,536 blocks x 11 doubles, 5.8 MB per device, held until the process
 * ends), so n_pairs is limited to 65,536 * 256 = 16,777,216 per call
 * (GROMOS_AMD_EINVAL beyond).  Every call on one device shares that array:
 * calls that overlap on a device (other streams, other threads) race.
 * Results equal gromos_amd_lj_crf_innerloop's: energies and virial bit for
 * bit, forces to FP64 atomic-order rounding. */
int gromos_amd_lj_crf_innerloop_noalloc(const double *pos, const double *charges,
                                        const uint32_t *iac, const uint32_t *pairs,
                                        int64_t n_pairs, const double *lj, int32_t n_types,
                                        const gromos_amd_crf *crf, int32_t boundary,
                                        const double box[3], double four_pi_eps_i,
                                        int32_t do_virial, double *forces, double *energies,
                                        double *virial, void *stream);

/* lj_crf_innerloop_cg_grouped[_novirial] (innerloops.rs:374-428).
 * groups: 4*n_groups u32 (ref_atom_i, ref_atom_j, start, end) (CGPairGroup, params.rs:145-154). */
int gromos_amd_lj_crf_innerloop_cg_grouped(const double *pos, const double *charges,
                                           const uint32_t *iac, const uint32_t *pairs,
                                           const uint32_t *groups, int64_t n_groups,
                                           const double *lj, int32_t n_types,
                                           const gromos_amd_crf *crf, int32_t boundary,
                                           const double box[3], double four_pi_eps_i,
                                           int32_t do_virial, double *forces, double *energies,
                                           double *virial, void *stream);

/* solvent_innerloop[_novirial] (innerloops.rs:679-733): sentinel pairs of
 * first atoms, atoms_per_solvent <= 4, tables from the first pair's molecule. */
int gromos_amd_solvent_innerloop(const double *pos, const double *charges, const uint32_t *iac,
                                 const uint32_t *pairs, int64_t n_pairs, const double *lj,
                                 int32_t n_types, const gromos_amd_crf *crf, int32_t boundary,
                                 const double box[3], int32_t atoms_per_solvent,
                                 double four_pi_eps_i, int32_t do_virial, double *forces,
                                 double *energies, double *virial, void *stream);

/* rf_excluded_interactions (excluded.rs:25-112).  excl_off/excl: CSR over all
 * atoms, sorted symmetric lists (Topology::exclusions). */
int gromos_amd_rf_excluded_interactions(const double *charges, const int32_t *excl_off,
                                        const int32_t *excl, int32_t n_atoms,
                                        int32_t num_solute_atoms, const double *pos,
                                        const gromos_amd_crf *crf, int32_t boundary,
                                        const double box[3], double four_pi_eps_i,
                                        double *forces, double *energies, double *virial,
                                        void *stream);

/* one_four_interaction_loop (excluded.rs:131-206).  o14_off/o14: CSR of
 * Topology::one_four_pairs (pairs with j > i are evaluated). */
int gromos_amd_one_four_interaction_loop(const int32_t *o14_off, const int32_t *o14,
                                         int32_t n_atoms, const double *pos,
                                         const double *charges, const uint32_t *iac,
                                         const double *lj, int32_t n_types,
                                         const gromos_amd_crf *crf, int32_t boundary,
                                         const double box[3], double four_pi_eps_i,
                                         double *forces, double *energies, double *virial,
                                         double coulomb_scaling, void *stream);

/* PerturbedLambdaParams (perturbed.rs:11-50), from_lambda (perturbed.rs:57-101). */
typedef struct {
    double a_lj_lambda_n, b_lj_lambda_n, a_crf_lambda_n, b_crf_lambda_n;
    double a_lj_lambda_n_1, b_lj_lambda_n_1, a_crf_lambda_n_1, b_crf_lambda_n_1;
    double a_ljs_lambda, b_ljs_lambda, a_ljs_lambda2, b_ljs_lambda2;
    double a_crfs_lambda, b_crfs_lambda, a_crfs_lambda2, b_crfs_lambda2;
    int32_t lambda_exp;
} gromos_amd_lambda;

int gromos_amd_lambda_from(double lj_lambda, double ljs_lambda, double crf_lambda,
                           double crfs_lambda, double lj_deriv, double ljs_deriv,
                           double crf_deriv, double crfs_deriv, int32_t n,
                           gromos_amd_lambda *out);

/* Per-atom perturbation info (PertAtomInfo, perturbed.rs:271-280), device arrays;
 * flag[i] != 0 marks a perturbed atom; alphas already include ALPHLJ / ALPHC. */
typedef struct {
    const uint8_t *flag;
    const int32_t *a_iac, *b_iac;
    const double *a_q, *b_q, *alpha_lj, *alpha_crf;
} gromos_amd_pert;

/* perturbed_pairlist_correction (perturbed.rs:309-386).  out_energies: 3 doubles
 * accumulated (delta_e_lj, delta_e_crf, dhdl); forces accumulate the delta. */
int gromos_amd_perturbed_pairlist_correction(const double *pos, const uint32_t *pairs,
                                             int64_t n_pairs, const gromos_amd_pert *pert,
                                             const double *charges, const uint32_t *iac,
                                             const double *lj, int32_t n_types,
                                             const gromos_amd_crf *crf,
                                             const gromos_amd_lambda *lp, int32_t boundary,
                                             const double box[3], double four_pi_eps_i,
                                             double *forces, double *out_energies, void *stream);

/* perturbed_self_energy_correction (perturbed.rs:395-425).  Perturbed atoms
 * are those with pert->flag set among n_atoms (the term reads their A / B
 * charges).  out: 2 doubles accumulated (delta_e_crf, dhdl). */
int gromos_amd_perturbed_self_energy_correction(const gromos_amd_pert *pert, int32_t n_atoms,
                                                const gromos_amd_crf *crf,
                                                const gromos_amd_lambda *lp, double four_pi_eps_i,
                                                double *out, void *stream);

/* perturbed_excluded_correction (perturbed.rs:431-511): solute atoms'
 * excluded pairs j > i with at least one perturbed atom; excl_off / excl as in
 * gromos_amd_rf_excluded_interactions.  out_energies: 3 doubles accumulated
 * (delta_e_lj = 0, delta_e_crf, dhdl); forces accumulate the delta. */
int gromos_amd_perturbed_excluded_correction(const int32_t *excl_off, const int32_t *excl,
                                             int32_t n_atoms, int32_t num_solute_atoms,
                                             const double *pos, const gromos_amd_pert *pert,
                                             const double *charges, const gromos_amd_crf *crf,
                                             const gromos_amd_lambda *lp, int32_t boundary,
                                             const double box[3], double four_pi_eps_i,
                                             double *forces, double *out_energies, void *stream);

/* perturbed_one_four_correction (perturbed.rs:517-605): 1-4 pairs j > i with
 * at least one perturbed atom, soft-core on the cs6/cs12 entries; o14_off /
 * o14 as in gromos_amd_one_four_interaction_loop.  out_energies: 3 doubles
 * accumulated (delta_e_lj, delta_e_crf, dhdl). */
int gromos_amd_perturbed_one_four_correction(const int32_t *o14_off, const int32_t *o14,
                                             int32_t n_atoms, const double *pos,
                                             const gromos_amd_pert *pert, const double *charges,
                                             const uint32_t *iac, const double *lj, int32_t n_types,
                                             const gromos_amd_crf *crf, const gromos_amd_lambda *lp,
                                             int32_t boundary, const double box[3],
                                             double four_pi_eps_i, double *forces,
                                             double *out_energies, void *stream);

/* perturbed_atom_pair_correction (perturbed.rs:624-707).  atom_pairs: 4 int32
 * per PERTATOMPAIR entry (i, j, a_type, b_type; types 0 = c6/c12, 1 =
 * cs6/cs12, b_type -1 = absent in B, PerturbedAtomPair); entries that are not
 * 1-4 pairs (o14_off / o14) are skipped.  out_energies: 3 doubles accumulated
 * (delta_e_lj, delta_e_crf, dhdl). */
int gromos_amd_perturbed_atom_pair_correction(const int32_t *atom_pairs, int64_t n_atom_pairs,
                                              const int32_t *o14_off, const int32_t *o14,
                                              int32_t n_atoms, const double *pos,
                                              const double *charges, const uint32_t *iac,
                                              const double *lj, int32_t n_types,
                                              const gromos_amd_crf *crf,
                                              const gromos_amd_lambda *lp, int32_t boundary,
                                              const double box[3], double four_pi_eps_i,
                                              double *forces, double *out_energies, void *stream);

/* perturbed_lj_crf_interaction (perturbed.rs:124-266) on n independent pairs
 * (device arrays): r = 3n distance vectors; prm = 8n values (a_c6, a_c12,
 * b_c6, b_c12, a_q, b_q, alpha_lj, alpha_crf) with a_q / b_q the pair's charge
 * products; out = 5n values assigned (force / r, e_lj, e_crf, de_lj, de_crf). */
int gromos_amd_perturbed_lj_crf_interaction(const double *r, const double *prm, int64_t n,
                                            const gromos_amd_lambda *lp, const gromos_amd_crf *crf,
                                            double four_pi_eps_i, double *out, void *stream);

/* Covalent terms of the solute (crates/gromos-forces/src/bonded): each index
 * row holds the atoms and a 0-based type; type tables as the reference stores
 * them (crates/gromos-io/src/topology.rs:373-710, units converted):
 *   bonds      i j t       bond_types     k_quartic k_harmonic r0        (3 per type)
 *   angles     i j k t     angle_types    k_cosine k_harmonic theta0/rad (3)
 *   impropers  i j k l t   improper_types k (kJ/mol/rad^2) q0/rad      (2)
 *   dihedrals  i j k l t   dihedral_types k cos(delta) m (integer 0..6)  (3)
 * A term whose type is beyond its table, or a dihedral of another
 * multiplicity, is skipped as the reference skips it.  ntf: the FORCE block's
 * NTF(1..4) switches (md.rs:331-335): bonds, angles, impropers, dihedrals. */
typedef struct {
    int32_t n_atoms;                /* the standalone plan: atoms of the position array */
    int32_t n_bonds, n_bond_types;
    const int32_t *bonds;
    const double *bond_types;
    int32_t n_angles, n_angle_types;
    const int32_t *angles;
    const double *angle_types;
    int32_t n_impropers, n_improper_types;
    const int32_t *impropers;
    const double *improper_types;
    int32_t n_dihedrals, n_dihedral_types;
    const int32_t *dihedrals;
    const double *dihedral_types;
    int32_t ntf[4];
} gromos_amd_bonded;

/* calculate_bonded_forces_ntf with quartic bonds (bonded/mod.rs:141-185,
 * forcefield.rs:386-400) as a plan: the terms are validated and laid out on the
 * device once (host inputs), then applied to device positions any number of
 * times.  Covalent forces only (3 n_atoms, assigned) and out[13] (device):
 * e_bond, e_angle, e_dihedral, e_improper, virial[9] (row a, column b).  FP64;
 * an atom's force is summed in the reference's order (term order within a kind,
 * kinds added bonds, angles, dihedrals, impropers).  The plan holds per-term
 * scratch: calls that overlap on one plan race. */
typedef struct gromos_amd_bonded_plan gromos_amd_bonded_plan;
int gromos_amd_bonded_plan_create(const gromos_amd_bonded *terms, gromos_amd_bonded_plan **out);
void gromos_amd_bonded_plan_destroy(gromos_amd_bonded_plan *plan);
/* terms kept per kind: bonds, angles, dihedrals, impropers */
int gromos_amd_bonded_plan_terms(const gromos_amd_bonded_plan *plan, int32_t counts[4]);
int gromos_amd_bonded_forces(const gromos_amd_bonded_plan *plan, const double *pos, double *forces,
                             double *out, void *stream);

/* ------------------------------------------------------------------ */
/* 2. device-resident systems                                          */
/* ------------------------------------------------------------------ */

/* Topology essentials of one job (host memory). */
typedef struct {
    int32_t n_atoms;
    int32_t n_solute_atoms;
    int32_t n_cg;
    int32_t n_solute_cg;
    const int32_t *cg_start;        /* n_cg + 1; contiguous charge groups */
    const double *charge;
    const int32_t *iac;
    const double *mass;
    int32_t n_types;
    const double *lj;               /* n_types^2 * 4 */
    const int32_t *excl_off, *excl; /* sorted symmetric exclusions */
    const int32_t *o14_off, *o14;   /* one_four_pairs as read */
    int32_t atoms_per_solvent;
    /* SETTLE template (constraints.rs:826-835): O-H and H-H lengths */
    double dist_oh, dist_hh;
    /* optional perturbation (NULL flag = none) */
    const uint8_t *pert_flag;
    const int32_t *pert_a_iac, *pert_b_iac;
    const double *pert_a_q, *pert_b_q, *pert_alpha_lj, *pert_alpha_crf;
    int32_t n_pert_atom_pairs;
    const int32_t *pert_atom_pairs; /* i, j, a_type, b_type (-1 = absent in B) */
    /* optional covalent terms of the solute atoms (NULL = none; n_atoms not read):
     * every force evaluation adds them, FP64 in both precisions, and e_pot
     * includes their energies (configuration.rs:197-207) */
    const gromos_amd_bonded *bonded;
    /* pressure groups of the solute (topology PRESSUREGROUPS): the last atom of
     * each, ascending, the last one n_solute_atoms - 1; every solvent molecule
     * is a group of its own.  Read only with a molecular virial (params.virial 2) */
    int32_t n_pressure_groups;
    const int32_t *pressure_group_last;
    /* SHAKE constraints (constraints.rs:213-570): the solvent molecule's template
     * (pairs of atom indices within the molecule, lengths; read with
     * params.solvent_shake), and the solute's (pairs of solute atoms, lengths;
     * NTC 2 / 3 as the caller selected them, in the reference's order) */
    int32_t n_solvent_constraints;
    const int32_t *solvent_constraint_ij;
    const double *solvent_constraint_len;
    int32_t n_solute_constraints;
    const int32_t *solute_constraint_ij;
    const double *solute_constraint_len;
} gromos_amd_topology;

/* Run parameters (IMD PAIRLIST / NONBONDED / STEP / PERTURBATION blocks). */
typedef struct {
    int32_t boundary;
    double box[3];
    double rcutp, rcutl, skin;
    int32_t nsnb;
    double epsrf, kappa;            /* CRF from rcutl, epsilon = 1 (md.rs:326-329) */
    double four_pi_eps_i;
    double dt;
    int32_t use_settle;
    int32_t do_virial;
    double lambda;
    int32_t nlam;
    /* one bath over the job's atoms: 0 = none (NVE), 1 = Berendsen weak
     * coupling (berendsen_thermostat.rs), 2 = Nose-Hoover chain of length nhc
     * (nosehoover_thermostat.rs; nhc 1 is the single Nose-Hoover thermostat);
     * reference temperature t_ref (K), coupling time tau_t (ps; < 0 = no
     * coupling, Berendsen: 0 = instantaneous) and dof_t degrees of freedom
     * (md.rs:1152-1172: 3N - constraints - NDFMIN) */
    int32_t thermostat;
    double t_ref, tau_t, dof_t;
    /* boundary 2 (triclinic, Triclinic in math.rs:75-141): the box vectors
     * a, b, c (9 doubles, lower triangular: a = (ax,0,0), b = (bx,by,0));
     * `box` is not read then */
    double box_matrix[9];
    int32_t nhc;                    /* thermostat 2: chain length, 1 .. GROMOS_AMD_NHC_MAX */
    /* centre-of-mass motion removal at the start of md steps (remove_com_motion.rs):
     * nticom 1 removes the translation at the job's step 0, 2 also the rotation;
     * nscm > 0 removes the translation every nscm steps, nscm < 0 the
     * translation and, without a periodic box, the rotation every |nscm| steps */
    int32_t nticom, nscm;
    /* pressure coupling at the end of md steps (md.rs:1116-1129, 1259-1277;
     * rectangular or triclinic boxes, every box vector scaled): barostat 1 =
     * isotropic Berendsen (berendsen_barostat.rs)
     * toward p_ref (kJ mol^-1 nm^-3) with compressibility (mol nm^3 kJ^-1, the
     * GROMOS input value) and tau_p (ps); the pressure (pressure_calculation.rs)
     * from the virial: 1 atomic, 2 molecular (prepare_virial, forcefield.rs:285-296,
     * 1128-1200), 0 none (the pressure stays 0) */
    int32_t barostat;
    double p_ref, compressibility, tau_p;
    int32_t virial;
    /* solvent_shake 1: the solvent molecules are constrained by SHAKE with the
     * topology's template instead of SETTLE; shake_tolerance: NTCP0 / NTCS0 (also
     * for the solute's constraints) */
    int32_t solvent_shake;
    double shake_tolerance;
    /* LINCS (constraints.rs:1168-1445, md.rs:1244-1257): solute_lincs > 0
     * constrains the solute's constraints (topology) by LINCS of that expansion
     * order (NTCP lincs, NTCP0) instead of SHAKE; solvent_lincs > 0 every solvent
     * molecule by the topology's template (NTCS lincs, NTCS0; at most 4 atoms and
     * 6 constraints) instead of SHAKE or SETTLE.  LINCS adds no constraint virial. */
    int32_t solute_lincs;
    int32_t solvent_lincs;
} gromos_amd_params;

#define GROMOS_AMD_NHC_MAX 16

enum gromos_amd_precision {
    GROMOS_AMD_FP64 = 0,   /* plain: FP64 throughout, reference-exact pair list */
    GROMOS_AMD_MIXED = 1   /* FP32 pair math, int64 fixed-point forces, FP64 sums/SETTLE */
};

typedef struct gromos_amd_system gromos_amd_system;

int gromos_amd_system_create(const gromos_amd_topology *topo, const gromos_amd_params *prm,
                             int32_t precision, gromos_amd_system **out);
void gromos_amd_system_destroy(gromos_amd_system *sys);
/* positions, velocities: host (n_atoms * 3) f64; vel may be NULL (zeros). */
int gromos_amd_system_set_state(gromos_amd_system *sys, const double *pos, const double *vel);
int gromos_amd_system_get_state(gromos_amd_system *sys, double *pos, double *vel);
/* Initial pair list before the step loop (md.rs:866-868); resets the rebuild counter. */
int gromos_amd_system_pairlist_update(gromos_amd_system *sys);
/* One Forcefield::apply on the current positions (rebuilds on its cadence).
 * Host outputs, any may be NULL: forces (3n), energies (e_lj, e_crf, dhdl), virial (9). */
int gromos_amd_system_forcefield_apply(gromos_amd_system *sys, double *forces, double *energies,
                                       double *virial);
/* Covalent energies of the last force evaluation (forcefield_apply, or the
 * last step of md): 4 per job, e_bond, e_angle, e_dihedral, e_improper. */
int gromos_amd_system_covalent_energies(gromos_amd_system *sys, double *out);
/* The periodic box edge lengths now (3 per job; a barostat changes them) and
 * the isotropic pressure of each job's last md step with a virial (1 per job;
 * kJ mol^-1 nm^-3); either output may be NULL. */
int gromos_amd_system_box(gromos_amd_system *sys, double *box, double *pressure);
/* nsteps of Forcefield -> LeapFrogVelocity -> LeapFrogPosition -> SETTLE ->
 * TemperatureCalculation (md.rs:1143-1286).  energies (optional, host): per step
 * 5 doubles (e_lj, e_crf, e_pot, e_kin, dhdl). */
int gromos_amd_system_md(gromos_amd_system *sys, int32_t nsteps, double *energies);
/* md with the covalent energies of every step beside (covalent: 4 per step and
 * job, e_bond, e_angle, e_dihedral, e_improper; either output may be NULL):
 * the .tre bond / angle / improper / dihedral columns (md.rs energy output). */
int gromos_amd_system_md_covalent(gromos_amd_system *sys, int32_t nsteps, double *energies, double *covalent);
/* Pair lists materialized as the reference's four atom-pair lists (test mode).
 * Pass NULL buffers to query counts (n_out[4]: solute_short, solute_long,
 * solvent_short, solvent_long); then call again with buffers of 2*n u32. */
int gromos_amd_system_pairlist_get(gromos_amd_system *sys, int64_t n_out[4],
                                   uint32_t *solute_short, uint32_t *solute_long,
                                   uint32_t *solvent_short, uint32_t *solvent_long);
/* Pair interactions evaluated (counted as in the reference: n_i * n_j per CG pair,
 * long pairs only when recomputed) since creation. */
int64_t gromos_amd_system_pair_count(const gromos_amd_system *sys);
/* Device timing of the dominant nonbonded kernel: hipEvents recorded around each
 * launch while enabled.  total_ms / launches since enable. */
int gromos_amd_system_kernel_timing(gromos_amd_system *sys, int32_t enable, double *total_ms,
                                    int64_t *launches);
/* Pair interactions the timed launches of the dominant kernel evaluated since
 * timing was enabled (the solvent-solvent pairs of the jobs that advanced:
 * the unit of that kernel's roofline).  No counterpart in the reference. */
int64_t gromos_amd_system_kernel_pairs(const gromos_amd_system *sys);
/* HBM bytes held by the system's device buffers. */
int64_t gromos_amd_system_device_bytes(const gromos_amd_system *sys);
/* Lanes per solvent molecule the autotuner chose for the pair kernel (4, 8, 16;
 * 0 before the first force evaluation).  GROMOS_AMD_VV_LANES in the environment
 * forces a value. */
int32_t gromos_amd_system_vv_lanes(const gromos_amd_system *sys);
/* md steps run as a replayed graph so far.  With GROMOS_AMD_GRAPHS=1 the plain
 * steps (no pair-list rebuild now or next, no outside state change, no
 * per-kernel timing) replay one captured graph each; by default every step is
 * launched kernel by kernel (faster on ROCm 7.2).  No counterpart in the reference. */
int64_t gromos_amd_system_graph_launches(const gromos_amd_system *sys);
/* Pair-list builds so far, and how many of them refilled their lists with exact
 * slots (a charge group outgrew the slot sized from its previous count).
 * GROMOS_AMD_TIGHT_SLOTS=1 in the environment sizes slots at exactly the
 * previous counts (a test hook: most builds then refill).  No counterpart in the
 * reference. */
int gromos_amd_system_rebuild_stats(const gromos_amd_system *sys, int64_t *rebuilds, int64_t *refills);

/* ------------------------------------------------------------------ */
/* 3. sessions: continuous batching of independent jobs on one device  */
/*    (no counterpart in the reference, which runs one system per      */
/*    process, md.rs, and steps replicas one after another, remd.rs)   */
/* ------------------------------------------------------------------ */

typedef struct gromos_amd_session gromos_amd_session;

enum gromos_amd_job_state {
    GROMOS_AMD_JOB_QUEUED = 0,   /* admitted, waiting for the next layout boundary */
    GROMOS_AMD_JOB_RUNNING = 1,
    GROMOS_AMD_JOB_DONE = 2,     /* all requested steps done */
    GROMOS_AMD_JOB_FAILED = 3    /* stopped (SETTLE failure or a diverged force); last_error() after job_query says why */
};

typedef struct {
    int32_t device;          /* HIP device ordinal */
    int64_t hbm_budget;      /* bytes of job footprint the session may reserve; 0 = 90 % of free memory */
    int32_t precision;       /* gromos_amd_precision, for every job */
    int32_t nsnb;            /* pair-list cadence shared by all jobs (params.nsnb must match) */
    int32_t admit_interval;  /* steps between batch layout changes, a multiple of nsnb; 0 = nsnb */
} gromos_amd_session_params;

typedef struct {
    int32_t queued, running, done, resident_jobs;
    int64_t device_bytes;    /* held by the resident batch */
    int64_t reserved_bytes;  /* footprints of queued and resident jobs */
    int64_t budget;
    int64_t clock;           /* session steps so far */
    int64_t layouts;         /* batch layout changes */
    double layout_ms;        /* host wall time spent in them */
    int64_t pair_count;      /* pair interactions evaluated, all jobs */
    double layout_concat_ms; /* part of layout_ms: concatenating the jobs' arrays on the host */
} gromos_amd_session_info;

int gromos_amd_session_create(const gromos_amd_session_params *p, gromos_amd_session **out);
void gromos_amd_session_destroy(gromos_amd_session *s);
/* Device bytes a job adds to a batch (estimate used for admission); < 0 on error. */
int64_t gromos_amd_job_footprint(const gromos_amd_topology *t, const gromos_amd_params *p);
/* Copy topology and state (host) into the session and queue the job for nsteps
 * md steps.  GROMOS_AMD_EBUSY when its footprint does not fit the remaining
 * budget now (retry after jobs finish), GROMOS_AMD_ENOMEM when it never will. */
int gromos_amd_job_admit(gromos_amd_session *s, const gromos_amd_topology *t, const gromos_amd_params *p,
                         const double *pos, const double *vel, int64_t nsteps, int64_t *job_id);
/* Advance the session clock by nsteps: every running job does one md step per
 * clock step until its budget is used; queued jobs join and finished jobs leave
 * at clock multiples of admit_interval (at once when nothing is running). */
int gromos_amd_session_step(gromos_amd_session *s, int32_t nsteps);
int gromos_amd_job_query(gromos_amd_session *s, int64_t job, int32_t *state, int64_t *steps_done);
/* Current positions / velocities (n_atoms * 3, host); either may be NULL. */
int gromos_amd_job_fetch(gromos_amd_session *s, int64_t job, double *pos, double *vel);
/* The job's box edge lengths now (3) and its last md step's pressure; either may be NULL. */
int gromos_amd_job_box(gromos_amd_session *s, int64_t job, double *box, double *pressure);
/* Energy records of steps [first, first + count): 5 doubles per step
 * (e_lj, e_crf, e_pot, e_kin, dhdl), as gromos_amd_system_md. */
int gromos_amd_job_energies(gromos_amd_session *s, int64_t job, int64_t first, int64_t count, double *out);
/* Remove a job (a resident one leaves at the next layout boundary). */
int gromos_amd_job_evict(gromos_amd_session *s, int64_t job);
/* Replica exchange support, resident jobs, between steps:
 * 0.5 sum m v^2 of the current velocities (Replica::temperature, replica.rs:343-347). */
int gromos_amd_jobs_kinetic_energy(gromos_amd_session *s, int32_t n, const int64_t *jobs, double *ekin);
/* v *= factor per job (Replica::rescale_velocities, replica.rs:355-371). */
int gromos_amd_jobs_scale_velocities(gromos_amd_session *s, int32_t n, const int64_t *jobs,
                                     const double *factors);
/* Exchange positions and velocities of two jobs with equal atom counts
 * (Replica::exchange_configuration, replica.rs:350-352); only at clock
 * multiples of nsnb, the next step rebuilds the pair lists. */
int gromos_amd_jobs_swap_configurations(gromos_amd_session *s, int64_t a, int64_t b);
int gromos_amd_session_get_info(gromos_amd_session *s, gromos_amd_session_info *info);
/* Device time of each session step (hipEvents) while enabled: enable = 1 clears
 * and starts, 0 stops, -1 leaves it; copies up to cap values, *n = how many. */
int gromos_amd_session_step_times(gromos_amd_session *s, int32_t enable, double *ms, int64_t cap, int64_t *n);
/* gromos_amd_system_kernel_timing over the session's batches (kept across
 * layout changes): hipEvents around each launch of the solvent pair kernel on
 * the batch's stream while enabled; total_ms, launches and the pair
 * interactions those launches evaluated since enable (any output may be NULL;
 * enable = 1 clears and starts, 0 stops, -1 leaves it). */
int gromos_amd_session_kernel_timing(gromos_amd_session *s, int32_t enable, double *total_ms,
                                     int64_t *launches, int64_t *pairs);

#ifdef __cplusplus
}
#endif
#endif
