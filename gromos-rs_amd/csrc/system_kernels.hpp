// Device pointer bundle and launchers of the device-resident kernels.
#pragma once
#include <hip/hip_ext.h>
#include "bonded.hpp"
#include "system.hpp"

namespace gamd {

struct SysPtrs {
    const JobDev* jobs;
    // original-order topology (global over the batch)
    const int32_t* cg_start;   // ncg, first (global) atom of each charge group
    const int32_t* cg_n;       // ncg, atoms per charge group
    const int32_t* cg_job;     // ncg
    const int32_t* atom_job;   // npad, -1 on padding
    const double* q;
    const int32_t* iac;
    const double* mass;
    const double* inv_mass;
    const double* lj;          // all jobs' tables, job.lj_off
    const int32_t* excl_off;   // npad + 1
    const int32_t* excl;
    const int32_t* o14_off;    // one_four_pairs as read (exclusion semantics)
    const int32_t* o14;
    const int32_t* o14s_off;   // symmetric 1-4 entries, bit 31 = owner is the lower atom
    const int32_t* o14s;
    // state (original order)
    const double* x_cur;
    double* x_new;
    const double* v_cur;
    double* v_new;
    double* f;
    // pair list
    double* ref;
    uint32_t* key_in;
    uint32_t* val_in;
    uint32_t* key_out;
    uint32_t* val_out;
    int32_t* cell_count;
    const int32_t* cell_start;
    int32_t* s_cg_orig;
    int32_t* s_cg_n;
    const int32_t* s_cg_scan;
    int32_t* s_cg_start;
    double* s_ref;
    int32_t* s2o;
    int32_t* s_atom_cg;
    int32_t* s_type;
    int64_t* vv_cnt;
    int64_t* gen_cnt;
    const int4* vv_chunk;      // nvv_chunk: solvent pair kernel chunks (first sorted group, end, job, 0)
    int32_t nvv_chunk;
    unsigned* vv_sched;        // the persistent far-pair kernel's counters (NXCD_SCHED, zero between launches)
    int32_t vv_persist;        // resident blocks of that kernel; 0: the one-wave-per-launch-slot form
    int32_t* step;             // md step index within the current md call (device counter)
    const float4* prec;        // 2 per solute atom: (a_q, b_q, alpha_lj, alpha_crf), (a/b soft crf_cut3i, a/b iac bits)
    int32_t* s_prec;           // sorted atom -> its prec record (perturbed atoms), -1 otherwise
    uint32_t* step_done;       // blocks of k_reduce finished (self-resetting)
    const int64_t* vv_off;
    const int64_t* gen_off;
    uint32_t* vv_ent;
    uint32_t* gen_ent;
    uint64_t* counters;        // [0] short pairs per step, [1] long pairs per rebuild,
                               // [2] vv entries, [3] gen entries, [4] groups in gsol
    int32_t* gsol;             // sorted solvent groups with general-list entries (any order)
    int64_t* pc_short;         // per sorted group: canonical short / long pair interactions
    int64_t* pc_long;
    int64_t* job_pairs;        // per job: short, long pair interactions of this build
    int64_t* vv_cap;           // single-pass fill: entries reserved per sorted group
    int64_t* gen_cap;
    int32_t* prev_vv;          // previous build's counts per original group
    int32_t* prev_gen;
    int32_t* overflow;         // set when a group had more entries than reserved
    // per step
    double4* xq;
    float4* xf;                // mixed: FP32 offset from the group's first atom, charge
    long long* fi_s;           // mixed: int64 fixed-point forces (sorted)
    long long* fi_long_s;      // mixed: twin-range cache
    int32_t nrow_blocks;       // row stride of the energy partials per slot
    double* f_s;
    double* f_long_s;
    double* f_x;
    double* rows;
    double* vv_e;              // per sorted group: solvent-solvent e_lj, e_crf, virial[9]
    double* vv_part;           // per job: VV_PARTS partial sums of vv_e and cg_ek
    double* cg_ek;             // per original group: E_kin, E_kin_new of its atoms, solvent excluded RF energy
    const int32_t* sol_atom;   // nsol: original index of every solute atom, job by job
    int32_t nsol;
    double* sol_e;             // 2 x nsol x ROW: per solute atom, short then long pairs
    int32_t vv_virial;         // vv_e carries the virial
    double* records;
    int32_t* settle_fail;      // per job: first failing step
    int32_t* force_fail;       // per job: first step with a non-finite or out-of-range force
    double* therm_scale;       // per job: the bath's velocity scale of the current step
    double* nh_zeta;           // per job: Nose-Hoover chain (GROMOS_AMD_NHC_MAX), as of the last step
    double* nh_pend;           // per job: the chain the current step's scale came from
    const int64_t* job_step0;  // per job: its md steps done before the current call
    double* press;             // per job: kinetic energy tensor (9), then the last pressure
    const int32_t* pg_start;   // pressure groups (atom index starts; JobDev pg_off / n_pg)
    int32_t* box_fail;         // per job: first step whose box outgrew the cell grid
    double* cg_cvir;           // per group: SETTLE's virial diagonal (pressure-coupled jobs)
    const int32_t* sh_ij;      // the solute's SHAKE constraints (atom pairs; JobDev sh_off / sh_n)
    const double* sh_len;
    uint8_t* sh_skip;          // two SHAKE skip flags per atom (the solute's sweeps)
    const double* lc_sdiag;    // LINCS of the solute's constraints (sh_off / sh_n): sdiag,
    const int32_t* lc_cpl;     // (start, count) of each constraint's coupled ones in
    const int32_t* lc_cidx;    //   lc_cidx (job-local constraint index, ascending) and
    const double* lc_coef;     //   lc_coef
    double* lc_work;           // per constraint: direction (3), two right-hand sides, solution
    // perturbation (original order; flag 0 = not perturbed)
    const uint8_t* pflag;
    const int32_t* pa_iac;
    const int32_t* pb_iac;
    const double* pa_q;
    const double* pb_q;
    const double* palj;
    const double* pacrf;
    const uint8_t* cg_pert;    // per original charge group: holds a perturbed atom
    const int32_t* papairs;    // PERTATOMPAIR: i, j, a_type, b_type (original indices)
    const int32_t* pap_off;    // npad + 1: entries of each atom
    const int32_t* pap;        // entry index | bit 31 when the atom is the entry's i
    int32_t* o2s;              // original atom -> sorted atom
    int32_t* vv_near;          // per sorted group: solvent partners at the start of its slot (near ones)
    long long* fi_near;        // mixed: the near solvent pairs' fixed-point forces (sorted), k_vv_nr
    double* vv_en;             // mixed: their per-group energy / virial record (11), k_vv_nr; else null
    BondedDev bd;              // covalent terms; bd.fb per solute atom (sol_atom order),
                               // nullptr when the batch has none
};

// nctl: words of the build's control block (counters) that k_cg_ref clears
void launch_cg_ref(const SysPtrs& p, int32_t ncg, int32_t nctl, hipStream_t s);
// stable counting sort of the groups by cell (cell_start scanned from cell_count)
// into s_cg_orig, s_cg_n, s_ref
void launch_sort_cg(const SysPtrs& p, int32_t ncg, int32_t ncell, hipStream_t s);
// mixed: also the FP32 offset copies xf
void launch_sorted_atoms(const SysPtrs& p, int32_t ncg, int32_t mixed, hipStream_t s);
// near_r2: solvent partners closer than this (squared, nm^2) go to the start of
// their group's slot, the others to its end (vv_near counts the former)
void launch_nbl(const SysPtrs& p, int32_t ncg, int32_t ncell, bool fill, double near_r2, hipStream_t s);
// per-job pair counts into job_pairs (cleared by the caller)
void launch_job_pairs(const SysPtrs& p, int32_t njobs, hipStream_t s);
// mode 0: capacities from the previous counts (kept by k_nbl) with slack; 1:
// capacities = counts; 2: the previous counts without slack (test hook)
void launch_caps(const SysPtrs& p, int32_t ncg, int32_t mode, hipStream_t s);
void launch_gather(const SysPtrs& p, int32_t npad, int32_t mixed, hipStream_t s);
// form: VV_LJ00 (LJ on the first-first atom pair only) | VV_NEUTRAL (neutral
// solvent molecules: the mixed kernel sums the RF polynomial per molecule pair);
// lanes: lanes per solvent molecule (4, 8 or 16)
constexpr int32_t VV_LJ00 = 1, VV_NEUTRAL = 2;
constexpr int32_t NXCD_SCHED = 9;   // the persistent far-pair kernel's counters: one per XCD, waves done
void launch_vv(const SysPtrs& p, int32_t ncg, int32_t aps, int32_t form, int32_t virial, int32_t mixed,
               int32_t lanes, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// mixed: the near solvent partners in FP64 (k_vv_nr) into fi_near / vv_en;
// independent of launch_vv's outputs
void launch_vv_near(const SysPtrs& p, int32_t aps, int32_t form, int32_t virial, hipStream_t s,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// ngsol: solvent groups listed in gsol by the last pair search (mixed short pass);
// part: GEN_ALL, or (mixed short pass) only the solvent or only the solute atoms
constexpr int32_t GEN_ALL = 0, GEN_SOLVENT = 1, GEN_SOLUTE = 2;
void launch_gen(const SysPtrs& p, int32_t npad, bool lng, int32_t virial, int32_t mixed, int32_t ngsol,
                hipStream_t s, int32_t part = GEN_ALL);
// solv_energy: also the solvent molecules' excluded RF energy (single points;
// md steps take it in the integrator)
// blocks (md steps, solv_energy 0): the nblocks atom blocks holding solute atoms
void launch_excl14(const SysPtrs& p, int32_t npad, int32_t solv_energy, const int32_t* blocks, int32_t nblocks,
                   hipStream_t s);
void launch_assemble(const SysPtrs& p, int32_t npad, int32_t with_long, int32_t mixed, hipStream_t s);
void launch_com(const SysPtrs& p, int32_t njobs, hipStream_t s);
void launch_press_ke(const SysPtrs& p, int32_t njobs, hipStream_t s);
void launch_baro(const SysPtrs& p, int32_t njobs, hipStream_t s);
// the solute's constraints by SHAKE or LINCS, a block per job (after the integrator)
void launch_shake_solute(const SysPtrs& p, int32_t njobs, hipStream_t s);
// SHAKE or LINCS of the solvent molecules, a lane per group (after the integrator)
void launch_solvent_constraints(const SysPtrs& p, int32_t ncg, hipStream_t s);
void launch_therm(const SysPtrs& p, int32_t njobs, hipStream_t s);
// leap-frog + Berendsen + SETTLE + E_kin (+ the next step's sorted copies) per group
void launch_integrate(const SysPtrs& p, int32_t ncg, int32_t with_long, int32_t write_xq,
                      int32_t mixed, hipStream_t s);
void launch_reduce(const SysPtrs& p, int32_t npad, int32_t with_long, int32_t with_kin, int32_t advance,
                   int32_t njobs, hipStream_t s);
void launch_materialize(const SysPtrs& p, int32_t ncg, bool fill, unsigned long long* counts, const int64_t* caps,
                        uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3, hipStream_t s);

}  // namespace gamd
