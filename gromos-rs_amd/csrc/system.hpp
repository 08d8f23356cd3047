// Device-resident MD jobs: data layout shared by the system kernels and the
// host driver.
//
// Layout in HBM (one "batch" = one or more independent jobs, concatenated):
//   * original atom space: job j owns [atom_off, atom_off + n_atoms); every job
//     starts on a BLOCK-aligned offset so that each block of a per-atom kernel
//     belongs to exactly one job (energies reduce per job without atomics);
//   * charge groups are contiguous atom ranges (topology.rs:1093-1110);
//   * sorted space: at every pair-list rebuild the charge groups of each job are
//     sorted by cell (stable radix sort on the cell key, so ties keep the
//     original order) and their atoms laid out contiguously inside the job's
//     range ("cell-sorted coordinates"); per step the positions are gathered
//     into xq (x, y, z, q as one 32-byte record per atom) in that order;
//   * neighbour lists are FULL charge-group lists in sorted space (every pair is
//     stored for both partners): one lane owns one atom and accumulates only
//     its own force, so no atomics are needed and the result does not depend on
//     scheduling.  Each entry carries the "canonical" bit (owner has the lower
//     original charge-group index): the pair is then evaluated in the
//     reference's orientation (lower group first) and only the canonical lane
//     adds energy and virial.
#pragma once
#include "device_common.hpp"

namespace gamd {

constexpr int BLOCK = 256;

// neighbour entry: partner sorted CG index | flags
constexpr uint32_t ENT_IDX = 0x0FFFFFFFu;
constexpr uint32_t ENT_VVLONG = 1u << 28;  // solvent-solvent long pair (expanded, per-atom PBC)
constexpr uint32_t ENT_SS = 1u << 29;      // both solute: exclusion / 1-4 check on short pairs
constexpr uint32_t ENT_LONG = 1u << 30;    // (RCUTP, RCUTL]
constexpr uint32_t ENT_CANON = 1u << 31;   // owner is the lower original charge group

struct JobDev {
    Box box;
    Crf crf;
    double fpepsi;
    double c2s, c2l;                       // (rc + skin)^2 short / long
    double sv_c6[16], sv_c12[16], sv_qq[16];  // solvent template tables (innerloops.rs:241-255)
    Lambda lp;
    int32_t n_types, lj_off;
    int32_t use_cg_grouped;                // forcefield.rs:271-283
    int32_t aps;                           // atoms per solvent molecule (<= 4)
    int32_t atom_off, n_atoms, n_solute_atoms;
    int32_t cg_off, n_cg, n_solute_cg;
    int32_t ncell[3], cell_off;
    int32_t twin;                          // rcutp < rcutl - 1e-10 (forcefield.rs:126)
    int32_t has_pert;
    int32_t stop;                          // the job advances only in md steps s < stop
                                           // of the current call (INT32_MAX: always)
    double dt;
    double m_o, m_h, dist_oh, dist_hh;     // SETTLE template
    int32_t sol_off;                       // first index of the job's solute atoms in sol_atom
    int32_t settle;                        // SETTLE the solvent molecules
    int32_t thermo;                        // 1 Berendsen (berendsen_thermostat.rs), 2 Nose-Hoover chain
    double t_ref, tau_t, dof_t;
    int32_t nhc;                           // Nose-Hoover chain length (nosehoover_thermostat.rs)
    int32_t nticom, nscm;                  // centre-of-mass motion removal (remove_com_motion.rs)
    int32_t baro, vtype;                   // Berendsen barostat; virial 0 none, 1 atomic, 2 molecular
    double p_ref, comp, tau_p;
    double cell_min;                       // the pair search's least cell width (half the list cutoff)
    double cgg_limit;                      // CG-grouped loops while half the shortest edge exceeds it
    int32_t pg_off, n_pg;                  // pressure groups [pg_start[pg_off + g], pg_start[pg_off + g + 1])
    int32_t shake_sv;                      // SHAKE the solvent molecules (template below)
    int32_t n_sc;                          // template constraints (<= 6), atom pairs within a molecule
    int32_t sc_ij[12];
    double sc_len[6];
    double sh_tol;                         // SHAKE tolerance
    int32_t sh_off, sh_n;                  // the solute's constraints in the batch arrays
    int32_t lincs_u;                       // LINCS order of the solute's constraints (0: SHAKE)
    int32_t lincs_sv;                      // LINCS order of the solvent template (0: none)
    int32_t lc_cpl[6];                     // template: bit c of lc_cpl[k] - constraints k, c share an atom
    double lc_sdiag[6];                    // template: 1 / sqrt(1/m_i + 1/m_j)
    double lc_coef[36];                    // template: coupling coefficients [k * 6 + c]
    int32_t bd_off;                        // first covalent term (batch order)
    int32_t bd_n[4];                       // terms per kind: bonds, angles, dihedrals, impropers
};

// energy partial rows: per kernel slot and block, 12 doubles
// (e_lj, e_crf, vir[9], extra)
constexpr int ROW = 12;
enum Slot { SLOT_VV = 0, SLOT_GEN = 1, SLOT_LONG = 2, SLOT_EXCL = 3, SLOT_KIN = 4, NSLOT = 5 };

// a force component at or beyond this (kJ mol^-1 nm^-1), or a non-finite one,
// marks its job as diverged (the fixed-point conversions saturate at twice it)
constexpr double FORCE_LIMIT = 33554432.0;   // 2^25
constexpr int NHC_MAX = 16;                   // Nose-Hoover chain slots per job (GROMOS_AMD_NHC_MAX)

// per step and job record: e_lj, e_crf, e_pot, e_kin, dhdl, vir[9],
// e_bond, e_angle, e_dihedral, e_improper
constexpr int REC = 18;
// words per job in the pair-list build's control block: short and long pair
// interactions, solvent-solvent list entries
constexpr int JOB_CTL = 3;

// parts per job of the solvent-solvent energy and E_kin records (two-stage
// reduction); PART_W columns: e_lj, e_crf, virial[9], E_kin, E_kin_new
constexpr int VV_CHUNK = 16;   // groups per chunk of the solvent pair kernels (one job each)
// mixed precision: solvent molecule pairs whose first atoms are closer than
// VV_NEAR_R (nm) at the list build are evaluated in FP64 (0.6 nm: the first
// two hydration shells, ~8 % of a 1.4 nm list; the largest FP32 force error
// over the 506,566 atoms of the 64-job FEP mix falls from 2.3e-5 to 6.7e-6 of
// the atom's force for ~6 % of the step time).  GROMOS_AMD_VV_NEAR_R in the
// environment sets another radius (0: all FP32).
constexpr double VV_NEAR_R = 0.6;
constexpr int VV_PARTS = 128;
constexpr int PART_W = 13;

}  // namespace gamd
