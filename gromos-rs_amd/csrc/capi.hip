// C ABI of gromos_amd (include/gromos_amd.h): argument checking, device memory,
// and the host-side step driver of the device-resident systems.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>

#include "explicit.hpp"
#include "host.hpp"

using namespace gamd;

static thread_local std::string g_err;

namespace gamd {
int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace gamd

extern "C" const char* gromos_amd_last_error(void) { return g_err.c_str(); }
extern "C" const char* gromos_amd_version(void) { return "gromos_amd 0.1 (gfx950)"; }

extern "C" int gromos_amd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int gromos_amd_crf_new(double cutoff, double epsilon, double rf_epsilon, double rf_kappa,
                                  gromos_amd_crf* o) {
    if (!o || !(cutoff > 0.0)) return fail(GROMOS_AMD_EINVAL, "crf_new: cutoff must be > 0");
    // CRFParameters::new (params.rs:78-97)
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
    return 0;
}

extern "C" int gromos_amd_lambda_from(double lj, double ljs, double crfl, double crfs, double ljd,
                                      double ljsd, double crfd, double crfsd, int32_t n,
                                      gromos_amd_lambda* o) {
    if (!o) return fail(GROMOS_AMD_EINVAL, "lambda_from: null output");
    // PerturbedLambdaParams::from_lambda (perturbed.rs:57-101); powi by squaring
    auto powi = [](double x, int k) {
        bool neg = k < 0;
        unsigned u = (unsigned)(neg ? -k : k);
        double r = 1.0;
        for (;;) {
            if (u & 1u) r *= x;
            u >>= 1;
            if (!u) break;
            x *= x;
        }
        return neg ? 1.0 / r : r;
    };
    o->a_lj_lambda_n = powi(1.0 - lj, n);
    o->a_crf_lambda_n = powi(1.0 - crfl, n);
    o->b_lj_lambda_n = powi(lj, n);
    o->b_crf_lambda_n = powi(crfl, n);
    o->a_lj_lambda_n_1 = powi(1.0 - lj, n - 1) * ljd;
    o->a_crf_lambda_n_1 = powi(1.0 - crfl, n - 1) * crfd;
    o->b_lj_lambda_n_1 = powi(lj, n - 1) * ljd;
    o->b_crf_lambda_n_1 = powi(crfl, n - 1) * crfd;
    o->a_ljs_lambda = (1.0 - ljs) * ljsd;
    o->a_crfs_lambda = (1.0 - crfs) * crfsd;
    o->b_ljs_lambda = ljs * ljsd;
    o->b_crfs_lambda = crfs * crfsd;
    o->a_ljs_lambda2 = (1.0 - ljs) * (1.0 - ljs);
    o->a_crfs_lambda2 = (1.0 - crfs) * (1.0 - crfs);
    o->b_ljs_lambda2 = ljs * ljs;
    o->b_crfs_lambda2 = crfs * crfs;
    o->lambda_exp = n;
    return 0;
}


// ---------------------------------------------------------------------------
// 1. pair-list kernels
// ---------------------------------------------------------------------------

// The per-block partial sums of one call (expl_blocks(n) rows of 11 doubles),
// from the device's stream-ordered memory pool, released on the same stream
// after the fold has read them.
struct PartRows {
    double* p = nullptr;
    hipStream_t st = nullptr;
    int alloc(int64_t n, hipStream_t s) {
        st = s;
        if (n <= 0) return 0;
        HIPCHK(hipMallocAsync((void**)&p, (size_t)expl_blocks(n) * EXPL_ROW * sizeof(double), st));
        return 0;
    }
    ~PartRows() {
        if (p) (void)hipFreeAsync(p, st);
    }
};
static int expl_common(ExplArgs& a, const double* pos, const double* q, const uint32_t* iac,
                       const double* lj, int32_t nt, const gromos_amd_crf* crf, int32_t boundary,
                       const double* box, double fpe, int32_t do_virial, double* forces,
                       double* energies, double* virial) {
    if (int rc = check_box(boundary, box)) return rc;
    if (!pos || !q || !iac || !lj || !crf || !forces || !energies || nt <= 0)
        return fail(GROMOS_AMD_EINVAL, "null argument or n_types <= 0");
    if (do_virial && !virial) return fail(GROMOS_AMD_EINVAL, "do_virial set but virial is null");
    a.pos = pos;
    a.q = q;
    a.iac = iac;
    a.lj = lj;
    a.n_types = nt;
    a.crf = to_crf(crf);
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.forces = forces;
    a.energies = energies;
    a.virial = do_virial ? virial : nullptr;
    return 0;
}

extern "C" int gromos_amd_lj_crf_innerloop(const double* pos, const double* charges, const uint32_t* iac,
                                           const uint32_t* pairs, int64_t n_pairs, const double* lj,
                                           int32_t n_types, const gromos_amd_crf* crf, int32_t boundary,
                                           const double box[3], double fpe, int32_t do_virial,
                                           double* forces, double* energies, double* virial, void* stream) {
    ExplArgs a{};
    if (int rc = expl_common(a, pos, charges, iac, lj, n_types, crf, boundary, box, fpe, do_virial, forces,
                             energies, virial))
        return rc;
    if (n_pairs < 0 || (n_pairs > 0 && !pairs)) return fail(GROMOS_AMD_EINVAL, "bad pair list");
    a.pairs = pairs;
    a.n = n_pairs;
    PartRows part;
    if (int rc = part.alloc(a.n, (hipStream_t)stream)) return rc;
    a.part = part.p;
    HIPCHK(launch_plain(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_lj_crf_innerloop_noalloc(const double* pos, const double* charges, const uint32_t* iac,
                                                   const uint32_t* pairs, int64_t n_pairs, const double* lj,
                                                   int32_t n_types, const gromos_amd_crf* crf, int32_t boundary,
                                                   const double box[3], double fpe, int32_t do_virial,
                                                   double* forces, double* energies, double* virial,
                                                   void* stream) {
    ExplArgs a{};
    if (int rc = expl_common(a, pos, charges, iac, lj, n_types, crf, boundary, box, fpe, do_virial, forces,
                             energies, virial))
        return rc;
    if (n_pairs < 0 || (n_pairs > 0 && !pairs)) return fail(GROMOS_AMD_EINVAL, "bad pair list");
    if (n_pairs > NOALLOC_MAX_PAIRS)
        return fail(GROMOS_AMD_EINVAL, "lj_crf_innerloop_noalloc: %lld pairs exceed the limit of %lld per call",
                    (long long)n_pairs, (long long)NOALLOC_MAX_PAIRS);
    a.pairs = pairs;
    a.n = n_pairs;
    HIPCHK(launch_plain_noalloc(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_lj_crf_innerloop_cg_grouped(const double* pos, const double* charges,
                                                      const uint32_t* iac, const uint32_t* pairs,
                                                      const uint32_t* groups, int64_t n_groups,
                                                      const double* lj, int32_t n_types,
                                                      const gromos_amd_crf* crf, int32_t boundary,
                                                      const double box[3], double fpe, int32_t do_virial,
                                                      double* forces, double* energies, double* virial,
                                                      void* stream) {
    ExplArgs a{};
    if (int rc = expl_common(a, pos, charges, iac, lj, n_types, crf, boundary, box, fpe, do_virial, forces,
                             energies, virial))
        return rc;
    if (n_groups < 0 || (n_groups > 0 && (!pairs || !groups))) return fail(GROMOS_AMD_EINVAL, "bad groups");
    a.pairs = pairs;
    a.groups = groups;
    a.n = n_groups;
    PartRows part;
    if (int rc = part.alloc(a.n, (hipStream_t)stream)) return rc;
    a.part = part.p;
    HIPCHK(launch_cg_grouped(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_solvent_innerloop(const double* pos, const double* charges, const uint32_t* iac,
                                            const uint32_t* pairs, int64_t n_pairs, const double* lj,
                                            int32_t n_types, const gromos_amd_crf* crf, int32_t boundary,
                                            const double box[3], int32_t aps, double fpe, int32_t do_virial,
                                            double* forces, double* energies, double* virial, void* stream) {
    ExplArgs a{};
    if (int rc = expl_common(a, pos, charges, iac, lj, n_types, crf, boundary, box, fpe, do_virial, forces,
                             energies, virial))
        return rc;
    // innerloops.rs:237-240 asserts n <= 4
    if (aps < 1 || aps > 4) return fail(GROMOS_AMD_EINVAL, "atoms_per_solvent must be 1..4 (got %d)", aps);
    if (n_pairs < 0 || (n_pairs > 0 && !pairs)) return fail(GROMOS_AMD_EINVAL, "bad pair list");
    a.pairs = pairs;
    a.n = n_pairs;
    a.aps = aps;
    PartRows part;
    if (int rc = part.alloc(a.n, (hipStream_t)stream)) return rc;
    a.part = part.p;
    HIPCHK(launch_solvent(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_rf_excluded_interactions(const double* charges, const int32_t* excl_off,
                                                   const int32_t* excl, int32_t n_atoms, int32_t n_solute,
                                                   const double* pos, const gromos_amd_crf* crf,
                                                   int32_t boundary, const double box[3], double fpe,
                                                   double* forces, double* energies, double* virial,
                                                   void* stream) {
    if (int rc = check_box(boundary, box)) return rc;
    if (!charges || !excl_off || !pos || !crf || !forces || !energies || !virial || n_atoms < 0)
        return fail(GROMOS_AMD_EINVAL, "null argument");
    ExclArgs a{};
    a.off = excl_off;
    a.idx = excl;
    a.n_atoms = n_atoms;
    a.n_solute = n_solute;
    a.pos = pos;
    a.q = charges;
    a.crf = to_crf(crf);
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.forces = forces;
    a.energies = energies;
    a.virial = virial;
    PartRows part;
    if (int rc = part.alloc(n_atoms, (hipStream_t)stream)) return rc;
    a.part = part.p;
    HIPCHK(launch_rf_excluded(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_one_four_interaction_loop(const int32_t* o14_off, const int32_t* o14, int32_t n_atoms,
                                                    const double* pos, const double* charges,
                                                    const uint32_t* iac, const double* lj, int32_t n_types,
                                                    const gromos_amd_crf* crf, int32_t boundary,
                                                    const double box[3], double fpe, double* forces,
                                                    double* energies, double* virial, double cs, void* stream) {
    if (int rc = check_box(boundary, box)) return rc;
    if (!o14_off || !pos || !charges || !iac || !lj || !crf || !forces || !energies || !virial || n_atoms < 0)
        return fail(GROMOS_AMD_EINVAL, "null argument");
    ExclArgs a{};
    a.off = o14_off;
    a.idx = o14;
    a.n_atoms = n_atoms;
    a.pos = pos;
    a.q = charges;
    a.iac = iac;
    a.lj = lj;
    a.n_types = n_types;
    a.crf = to_crf(crf);
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.coulomb_scaling = cs;
    a.forces = forces;
    a.energies = energies;
    a.virial = virial;
    PartRows part;
    if (int rc = part.alloc(n_atoms, (hipStream_t)stream)) return rc;
    a.part = part.p;
    HIPCHK(launch_one_four(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_perturbed_pairlist_correction(const double* pos, const uint32_t* pairs, int64_t n_pairs,
                                                        const gromos_amd_pert* pert, const double* charges,
                                                        const uint32_t* iac, const double* lj, int32_t n_types,
                                                        const gromos_amd_crf* crf, const gromos_amd_lambda* lp,
                                                        int32_t boundary, const double box[3], double fpe,
                                                        double* forces, double* out, void* stream) {
    if (int rc = check_box(boundary, box)) return rc;
    if (!pos || !pert || !pert->flag || !charges || !iac || !lj || !crf || !lp || !forces || !out)
        return fail(GROMOS_AMD_EINVAL, "null argument");
    PertArgs a{};
    a.pos = pos;
    a.pairs = pairs;
    a.n = n_pairs;
    a.flag = pert->flag;
    a.a_iac = pert->a_iac;
    a.b_iac = pert->b_iac;
    a.a_q = pert->a_q;
    a.b_q = pert->b_q;
    a.alpha_lj = pert->alpha_lj;
    a.alpha_crf = pert->alpha_crf;
    a.q = charges;
    a.iac = iac;
    a.lj = lj;
    a.n_types = n_types;
    a.crf = to_crf(crf);
    static_assert(sizeof(Lambda) == sizeof(gromos_amd_lambda) || sizeof(Lambda) == sizeof(gromos_amd_lambda) + 4,
                  "lambda layout");
    std::memcpy(&a.lp, lp, sizeof(gromos_amd_lambda));
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.forces = forces;
    a.out = out;
    HIPCHK(launch_pert_pairs(a, (hipStream_t)stream));
    return 0;
}

// common part of the standalone correction passes: the per-atom perturbation
// arrays, CRF, lambda, box and outputs
static int pert_common(PertArgs& a, const gromos_amd_pert* pert, const double* charges,
                       const gromos_amd_crf* crf, const gromos_amd_lambda* lp, int32_t boundary,
                       const double* box, double fpe, double* forces, double* out, bool need_forces) {
    if (int rc = check_box(boundary, box)) return rc;
    if (!pert || !pert->flag || !pert->a_q || !pert->b_q || !pert->alpha_lj || !pert->alpha_crf || !charges ||
        !crf || !lp || !out || (need_forces && !forces))
        return fail(GROMOS_AMD_EINVAL, "null argument");
    a.flag = pert->flag;
    a.a_iac = pert->a_iac;
    a.b_iac = pert->b_iac;
    a.a_q = pert->a_q;
    a.b_q = pert->b_q;
    a.alpha_lj = pert->alpha_lj;
    a.alpha_crf = pert->alpha_crf;
    a.q = charges;
    a.crf = to_crf(crf);
    std::memcpy(&a.lp, lp, sizeof(gromos_amd_lambda));
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.forces = forces;
    a.out = out;
    return 0;
}

extern "C" int gromos_amd_perturbed_self_energy_correction(const gromos_amd_pert* pert, int32_t n_atoms,
                                                           const gromos_amd_crf* crf, const gromos_amd_lambda* lp,
                                                           double fpe, double* out, void* stream) {
    static const double no_box[3] = {0.0, 0.0, 0.0};
    static const double no_charge = 0.0;  // the self term reads only the A / B charges
    PertArgs a{};
    if (int rc = pert_common(a, pert, &no_charge, crf, lp, 0, no_box, fpe, nullptr, out, false)) return rc;
    if (n_atoms < 0) return fail(GROMOS_AMD_EINVAL, "n_atoms < 0");
    a.n_atoms = n_atoms;
    HIPCHK(launch_pert_self(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_perturbed_excluded_correction(const int32_t* excl_off, const int32_t* excl,
                                                        int32_t n_atoms, int32_t n_solute, const double* pos,
                                                        const gromos_amd_pert* pert, const double* charges,
                                                        const gromos_amd_crf* crf, const gromos_amd_lambda* lp,
                                                        int32_t boundary, const double box[3], double fpe,
                                                        double* forces, double* out, void* stream) {
    PertArgs a{};
    if (int rc = pert_common(a, pert, charges, crf, lp, boundary, box, fpe, forces, out, true)) return rc;
    if (!pos || !excl_off || n_atoms < 0 || n_solute < 0 || n_solute > n_atoms)
        return fail(GROMOS_AMD_EINVAL, "bad exclusion arguments (n_atoms %d, n_solute %d)", n_atoms, n_solute);
    a.pos = pos;
    a.off = excl_off;
    a.idx = excl;
    a.n_atoms = n_atoms;
    a.n_solute = n_solute;
    HIPCHK(launch_pert_excluded(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_perturbed_one_four_correction(const int32_t* o14_off, const int32_t* o14,
                                                        int32_t n_atoms, const double* pos,
                                                        const gromos_amd_pert* pert, const double* charges,
                                                        const uint32_t* iac, const double* lj, int32_t n_types,
                                                        const gromos_amd_crf* crf, const gromos_amd_lambda* lp,
                                                        int32_t boundary, const double box[3], double fpe,
                                                        double* forces, double* out, void* stream) {
    PertArgs a{};
    if (int rc = pert_common(a, pert, charges, crf, lp, boundary, box, fpe, forces, out, true)) return rc;
    if (!pos || !o14_off || !iac || !lj || !pert->a_iac || !pert->b_iac || n_atoms < 0 || n_types < 1)
        return fail(GROMOS_AMD_EINVAL, "bad 1-4 arguments");
    a.pos = pos;
    a.off = o14_off;
    a.idx = o14;
    a.n_atoms = n_atoms;
    a.iac = iac;
    a.lj = lj;
    a.n_types = n_types;
    HIPCHK(launch_pert_one_four(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_perturbed_atom_pair_correction(const int32_t* atom_pairs, int64_t n_atom_pairs,
                                                         const int32_t* o14_off, const int32_t* o14,
                                                         int32_t n_atoms, const double* pos, const double* charges,
                                                         const uint32_t* iac, const double* lj, int32_t n_types,
                                                         const gromos_amd_crf* crf, const gromos_amd_lambda* lp,
                                                         int32_t boundary, const double box[3], double fpe,
                                                         double* forces, double* out, void* stream) {
    if (int rc = check_box(boundary, box)) return rc;
    if (n_atom_pairs < 0 || (n_atom_pairs > 0 && !atom_pairs) || !o14_off || !pos || !charges || !iac || !lj ||
        !crf || !lp || !forces || !out || n_atoms < 0 || n_types < 1)
        return fail(GROMOS_AMD_EINVAL, "null or bad argument");
    PertArgs a{};
    a.atom_pairs = atom_pairs;
    a.n = n_atom_pairs;
    a.off = o14_off;
    a.idx = o14;
    a.n_atoms = n_atoms;
    a.pos = pos;
    a.q = charges;
    a.iac = iac;
    a.lj = lj;
    a.n_types = n_types;
    a.crf = to_crf(crf);
    std::memcpy(&a.lp, lp, sizeof(gromos_amd_lambda));
    a.box = make_box(boundary, box);
    a.fpepsi = fpe;
    a.forces = forces;
    a.out = out;
    HIPCHK(launch_pert_atom_pairs(a, (hipStream_t)stream));
    return 0;
}

extern "C" int gromos_amd_perturbed_lj_crf_interaction(const double* r, const double* prm, int64_t n,
                                                       const gromos_amd_lambda* lp, const gromos_amd_crf* crf,
                                                       double fpe, double* out, void* stream) {
    if (n < 0 || (n > 0 && (!r || !prm || !out)) || !lp || !crf) return fail(GROMOS_AMD_EINVAL, "null or bad argument");
    SoftPairArgs a{};
    a.r = r;
    a.prm = prm;
    a.n = n;
    std::memcpy(&a.lp, lp, sizeof(gromos_amd_lambda));
    a.crf = to_crf(crf);
    a.fpepsi = fpe;
    a.out = out;
    HIPCHK(launch_soft_pairs(a, (hipStream_t)stream));
    return 0;
}

// ---------------------------------------------------------------------------
// 2. device-resident systems
// ---------------------------------------------------------------------------
static int create_batch(const gromos_amd_topology* topos, const gromos_amd_params* prms, int njobs,
                        int32_t precision, gromos_amd_system** out) {
    if (!out || njobs <= 0 || !topos || !prms) return fail(GROMOS_AMD_EINVAL, "create: bad arguments");
    if (precision != GROMOS_AMD_FP64 && precision != GROMOS_AMD_MIXED)
        return fail(GROMOS_AMD_EINVAL, "precision must be 0 (FP64) or 1 (mixed)");
    *out = nullptr;
    std::vector<JobImage> imgs(njobs);
    std::vector<const JobImage*> ptr(njobs);
    for (int j = 0; j < njobs; ++j) {
        if (int rc = make_image(topos[j], prms[j], j, imgs[j])) return rc;
        ptr[j] = &imgs[j];
    }
    auto S = std::make_unique<gromos_amd_system>();
    if (int rc = layout_batch(S.get(), ptr, precision)) return rc;
    *out = S.release();
    return 0;
}

extern "C" int gromos_amd_system_create(const gromos_amd_topology* topo, const gromos_amd_params* prm,
                                        int32_t precision, gromos_amd_system** out) {
    return create_batch(topo, prm, 1, precision, out);
}

extern "C" void gromos_amd_system_destroy(gromos_amd_system* S) { delete S; }

extern "C" int64_t gromos_amd_system_device_bytes(const gromos_amd_system* S) { return S ? S->bytes : 0; }
extern "C" int32_t gromos_amd_system_vv_lanes(const gromos_amd_system* S) { return S ? S->vv_lanes : 0; }
extern "C" int64_t gromos_amd_system_graph_launches(const gromos_amd_system* S) { return S ? S->graph_launches : 0; }
extern "C" int gromos_amd_system_rebuild_stats(const gromos_amd_system* S, int64_t* rebuilds, int64_t* refills) {
    if (!S || !rebuilds || !refills) return fail(GROMOS_AMD_EINVAL, "rebuild_stats: null argument");
    *rebuilds = S->rebuilds;
    *refills = S->refills;
    return 0;
}
extern "C" int64_t gromos_amd_system_pair_count(const gromos_amd_system* S) { return S ? S->pair_count : 0; }

extern "C" int gromos_amd_system_set_state(gromos_amd_system* S, const double* pos, const double* vel) {
    if (!S || !pos) return fail(GROMOS_AMD_EINVAL, "set_state: null argument");
    size_t hoff = 0;
    for (int j = 0; j < S->njobs; ++j) {
        size_t off = (size_t)S->jobs[j].atom_off * 3, n3 = (size_t)S->job_n_atoms[j] * 3;
        HIPCHK(hipMemcpyAsync(S->d_x[S->cur] + off, pos + hoff, n3 * sizeof(double), hipMemcpyHostToDevice, S->stream));
        if (vel) HIPCHK(hipMemcpyAsync(S->d_v[S->cur] + off, vel + hoff, n3 * sizeof(double), hipMemcpyHostToDevice, S->stream));
        else HIPCHK(hipMemsetAsync(S->d_v[S->cur] + off, 0, n3 * sizeof(double), S->stream));
        hoff += n3;
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    S->built = false;
    S->long_computed = false;
    S->therm_dirty = true;
    S->xq_valid = false;
    return 0;
}

extern "C" int gromos_amd_system_get_state(gromos_amd_system* S, double* pos, double* vel) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "get_state: null system");
    size_t hoff = 0;
    for (int j = 0; j < S->njobs; ++j) {
        size_t off = (size_t)S->jobs[j].atom_off * 3, n3 = (size_t)S->job_n_atoms[j] * 3;
        if (pos) HIPCHK(hipMemcpyAsync(pos + hoff, S->d_x[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        if (vel) HIPCHK(hipMemcpyAsync(vel + hoff, S->d_v[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        hoff += n3;
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    return 0;
}

extern "C" int gromos_amd_system_pairlist_update(gromos_amd_system* S) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    return build_pairlist(S);
}

extern "C" int gromos_amd_system_forcefield_apply(gromos_amd_system* S, double* forces, double* energies,
                                                  double* virial) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    if (int rc = S->reserve(S->d_records, (size_t)S->njobs * REC, true)) return rc;
    if (int rc = forcefield(S, true)) return rc;
    SysPtrs p = S->ptrs();
    launch_reduce(p, S->npad, S->twin ? 1 : 0, 0, 0, S->njobs, S->stream);
    HIPCHK(hipGetLastError());
    S->last_rec_row = 0;
    std::vector<double> rec((size_t)S->njobs * REC);
    HIPCHK(hipMemcpyAsync(rec.data(), S->d_records, rec.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    size_t hoff = 0;
    for (int j = 0; j < S->njobs; ++j) {
        size_t off = (size_t)S->jobs[j].atom_off * 3, n3 = (size_t)S->job_n_atoms[j] * 3;
        if (forces) HIPCHK(hipMemcpyAsync(forces + hoff, S->d_f + off, n3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        hoff += n3;
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    for (int j = 0; j < S->njobs; ++j) {
        const double* r = rec.data() + (size_t)j * REC;
        if (energies) {
            energies[3 * j] = r[0];
            energies[3 * j + 1] = r[1];
            energies[3 * j + 2] = r[4];
        }
        if (virial) std::memcpy(virial + 9 * j, r + 5, 9 * sizeof(double));
    }
    return S->collect_timing();
}

extern "C" int gromos_amd_system_covalent_energies(gromos_amd_system* S, double* out) {
    if (!S || !out) return fail(GROMOS_AMD_EINVAL, "covalent_energies: bad arguments");
    if (S->last_rec_row < 0) return fail(GROMOS_AMD_EINVAL, "covalent_energies: no force evaluation yet");
    std::vector<double> rec((size_t)S->njobs * REC);
    HIPCHK(hipMemcpyAsync(rec.data(), S->d_records + (size_t)S->last_rec_row * S->njobs * REC,
                          rec.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    for (int j = 0; j < S->njobs; ++j)
        for (int k = 0; k < 4; ++k) out[4 * j + k] = rec[(size_t)j * REC + 14 + k];
    return 0;
}

extern "C" int gromos_amd_system_box(gromos_amd_system* S, double* box, double* pressure) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    for (int32_t b = 0; b < S->njobs; ++b) {
        if (box)
            for (int a = 0; a < 3; ++a) box[3 * b + a] = S->jobs[b].box.L[a];
        if (pressure) pressure[b] = S->host_pressure[b];
    }
    return 0;
}

extern "C" int gromos_amd_system_md_covalent(gromos_amd_system* S, int32_t nsteps, double* energies,
                                             double* covalent) {
    if (!S || nsteps < 0) return fail(GROMOS_AMD_EINVAL, "md: bad arguments");
    if (nsteps == 0) return 0;
    std::vector<double> rec;
    if (int rc = md_steps(S, nsteps, (energies || covalent) ? &rec : nullptr)) return rc;
    for (int32_t s = 0; s < nsteps && (energies || covalent); ++s)
        for (int j = 0; j < S->njobs; ++j) {
            const size_t row = (size_t)s * S->njobs + j;
            const double* r = rec.data() + row * REC;
            if (energies)
                for (int c = 0; c < 5; ++c) energies[row * 5 + c] = r[c];
            if (covalent)
                for (int c = 0; c < 4; ++c) covalent[row * 4 + c] = r[14 + c];
        }
    return 0;
}

extern "C" int gromos_amd_system_md(gromos_amd_system* S, int32_t nsteps, double* energies) {
    return gromos_amd_system_md_covalent(S, nsteps, energies, nullptr);
}

extern "C" int gromos_amd_system_pairlist_get(gromos_amd_system* S, int64_t n_out[4], uint32_t* ss, uint32_t* sl,
                                              uint32_t* vs, uint32_t* vl) {
    if (!S || !n_out) return fail(GROMOS_AMD_EINVAL, "pairlist_get: bad arguments");
    if (S->njobs != 1) return fail(GROMOS_AMD_EINVAL, "pairlist_get: single-job systems only");
    if (!S->built) {
        if (int rc = build_pairlist(S)) return rc;
    }
    unsigned long long* d_cnt = nullptr;
    HIPCHK(hipMalloc(&d_cnt, 8 * sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(d_cnt, 0, 4 * sizeof(unsigned long long), S->stream));
    int64_t caps[4] = {0, 0, 0, 0};
    bool fill = ss && sl && vs && vl;
    uint32_t* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
    if (fill) {
        for (int k = 0; k < 4; ++k) caps[k] = n_out[k];
        for (int k = 0; k < 4; ++k)
            if (n_out[k] > 0 && hipMalloc(&bufs[k], (size_t)n_out[k] * 2 * sizeof(uint32_t)) != hipSuccess) {
                for (int m = 0; m < k; ++m) (void)hipFree(bufs[m]);
                (void)hipFree(d_cnt);
                return fail(GROMOS_AMD_ENOMEM, "pairlist_get: allocation failed");
            }
    }
    SysPtrs p = S->ptrs();
    int64_t* d_caps = (int64_t*)(d_cnt + 4);
    HIPCHK(hipMemcpyAsync(d_caps, caps, sizeof(caps), hipMemcpyHostToDevice, S->stream));
    launch_materialize(p, S->ncg, fill, d_cnt, d_caps, bufs[0], bufs[1], bufs[2], bufs[3], S->stream);
    unsigned long long cnt[4];
    HIPCHK(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    uint32_t* hout[4] = {ss, sl, vs, vl};
    int rc = 0;
    for (int k = 0; k < 4; ++k) {
        if (fill) {
            if ((int64_t)cnt[k] != n_out[k]) rc = fail(GROMOS_AMD_EINVAL, "pairlist_get: list %d changed size", k);
            else if (cnt[k]) (void)hipMemcpy(hout[k], bufs[k], cnt[k] * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost);
            (void)hipFree(bufs[k]);
        }
        n_out[k] = (int64_t)cnt[k];
    }
    (void)hipFree(d_cnt);
    return rc;
}

extern "C" int gromos_amd_system_kernel_timing(gromos_amd_system* S, int32_t enable, double* total_ms,
                                               int64_t* launches) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    if (int rc = S->collect_timing()) return rc;
    if (total_ms) *total_ms = S->timed_ms;
    if (launches) *launches = S->timed_launches;
    if (enable >= 0) {
        bool on = enable != 0;
        if (on && !S->timing) {
            S->timed_ms = 0.0;
            S->timed_launches = 0;
            S->timed_pairs = 0;
        }
        S->timing = on;
    }
    return 0;
}

extern "C" int64_t gromos_amd_system_kernel_pairs(const gromos_amd_system* S) { return S ? S->timed_pairs : -1; }
