// Argument blocks of the pair-list kernels (explicit.hip).
#pragma once
#include "device_common.hpp"

namespace gamd {

struct ExplArgs {
    const double* pos;
    const double* q;
    const uint32_t* iac;
    const uint32_t* pairs;   // 2 per pair (plain, solvent) or the flat pair array (cg_grouped)
    const uint32_t* groups;  // cg_grouped only
    int64_t n;               // pairs, sentinels or groups
    const double* lj;
    int32_t n_types;
    int32_t aps;
    Crf crf;
    Box box;
    double fpepsi;
    double* forces;
    double* energies;
    double* virial;          // null: no virial
    double* part;            // scratch: 11 partial sums per block (blocks_for(n) rows)
};

struct ExclArgs {
    const int32_t* off;
    const int32_t* idx;
    int32_t n_atoms;
    int32_t n_solute;
    const double* pos;
    const double* q;
    const uint32_t* iac;
    const double* lj;
    int32_t n_types;
    Crf crf;
    Box box;
    double fpepsi;
    double coulomb_scaling;
    double* forces;
    double* energies;
    double* virial;
    double* part;            // scratch: 11 partial sums per block
};

constexpr int EXPL_BLOCK = 256;
constexpr int EXPL_ROW = 11;   // e_lj, e_crf, virial[9]
inline int64_t expl_blocks(int64_t n) { return (n + EXPL_BLOCK - 1) / EXPL_BLOCK; }
// Largest pair count of one gromos_amd_lj_crf_innerloop_noalloc call: its
// partial sums live in a static device array of NOALLOC_BLOCKS rows
// (65,536 x 11 doubles, 5.8 MB per device).
constexpr int64_t NOALLOC_BLOCKS = 65536;
constexpr int64_t NOALLOC_MAX_PAIRS = NOALLOC_BLOCKS * EXPL_BLOCK;

struct PertArgs {
    const double* pos;
    const uint32_t* pairs;
    int64_t n;
    const uint8_t* flag;
    const int32_t* a_iac;
    const int32_t* b_iac;
    const double* a_q;
    const double* b_q;
    const double* alpha_lj;
    const double* alpha_crf;
    const double* q;
    const uint32_t* iac;
    const double* lj;
    int32_t n_types;
    Crf crf;
    Lambda lp;
    Box box;
    double fpepsi;
    double* forces;
    double* out;  // delta_e_lj, delta_e_crf, dhdl (self pass: delta_e, dhdl)
    // excluded / 1-4 passes: CSR over atoms; atom-pair pass: entries and 1-4 CSR
    const int32_t* off;
    const int32_t* idx;
    int32_t n_atoms;
    int32_t n_solute;
    const int32_t* atom_pairs;  // i, j, a_type, b_type per entry
};

// perturbed_lj_crf_interaction over n independent pairs
struct SoftPairArgs {
    const double* r;    // 3 per pair
    const double* prm;  // a_c6, a_c12, b_c6, b_c12, a_q, b_q, alpha_lj, alpha_crf per pair
    int64_t n;
    Lambda lp;
    Crf crf;
    double fpepsi;
    double* out;        // f, e_lj, e_crf, de_lj, de_crf per pair
};

hipError_t launch_plain(const ExplArgs& a, hipStream_t s);
// launch_plain with the partial sums in the library's static device array
// instead of a.part (no allocation; calls that overlap on one device race)
hipError_t launch_plain_noalloc(ExplArgs a, hipStream_t s);
hipError_t launch_cg_grouped(const ExplArgs& a, hipStream_t s);
hipError_t launch_solvent(const ExplArgs& a, hipStream_t s);
hipError_t launch_rf_excluded(const ExclArgs& a, hipStream_t s);
hipError_t launch_one_four(const ExclArgs& a, hipStream_t s);
hipError_t launch_pert_pairs(const PertArgs& a, hipStream_t s);
hipError_t launch_pert_self(const PertArgs& a, hipStream_t s);
hipError_t launch_pert_excluded(const PertArgs& a, hipStream_t s);
hipError_t launch_pert_one_four(const PertArgs& a, hipStream_t s);
hipError_t launch_pert_atom_pairs(const PertArgs& a, hipStream_t s);
hipError_t launch_soft_pairs(const SoftPairArgs& a, hipStream_t s);

}  // namespace gamd
