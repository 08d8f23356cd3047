// C ABI of gromos_amd (include/gromos_amd.h): argument checking, device memory,
// and the host-side step driver of the device-resident systems.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gromos_amd.h"
#include "explicit.hpp"
#include "system_kernels.hpp"

using namespace gamd;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) return fail(GROMOS_AMD_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

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

static Box make_box(int32_t boundary, const double* box) {
    Box b{};
    b.kind = boundary == 1 ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        b.L[a] = b.kind ? box[a] : 0.0;
        b.invL[a] = b.kind ? 1.0 / box[a] : 0.0;   // DVec3::recip
    }
    return b;
}

static int check_box(int32_t boundary, const double* box) {
    if (boundary != 0 && boundary != 1) return fail(GROMOS_AMD_EINVAL, "boundary must be 0 (vacuum) or 1 (rectangular)");
    if (boundary == 1) {
        if (!box) return fail(GROMOS_AMD_EINVAL, "rectangular boundary needs a box");
        for (int a = 0; a < 3; ++a)
            if (!(box[a] > 0.0)) return fail(GROMOS_AMD_EINVAL, "box edge %d must be > 0", a);
    }
    return 0;
}

static Crf to_crf(const gromos_amd_crf* c) { return Crf{c->crf_cut, c->crf_2cut3i, c->crf_cut3i, c->cutoff_sq}; }

// ---------------------------------------------------------------------------
// 1. pair-list kernels
// ---------------------------------------------------------------------------
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
    HIPCHK(launch_plain(a, (hipStream_t)stream));
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

// ---------------------------------------------------------------------------
// 2. device-resident systems
// ---------------------------------------------------------------------------
struct gromos_amd_system {
    int precision = 0;
    int njobs = 0;
    int32_t npad = 0, ncg = 0, ncell = 0, nmol = 0;
    std::vector<JobDev> jobs;
    std::vector<int32_t> job_n_atoms;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    hipStream_t stream = nullptr;

    // device arrays
    JobDev* d_jobs = nullptr;
    int32_t *d_cg_start = nullptr, *d_cg_n = nullptr, *d_cg_job = nullptr, *d_atom_job = nullptr, *d_iac = nullptr;
    double *d_q = nullptr, *d_mass = nullptr, *d_inv_mass = nullptr, *d_lj = nullptr;
    int32_t *d_excl_off = nullptr, *d_excl = nullptr, *d_o14_off = nullptr, *d_o14 = nullptr;
    int32_t *d_o14s_off = nullptr, *d_o14s = nullptr, *d_mol_first = nullptr;
    double *d_x[2] = {nullptr, nullptr}, *d_v[2] = {nullptr, nullptr}, *d_f = nullptr;
    double *d_ref = nullptr, *d_s_ref = nullptr;
    uint32_t *d_key_in = nullptr, *d_val_in = nullptr, *d_key_out = nullptr, *d_val_out = nullptr;
    int32_t *d_cell_count = nullptr, *d_cell_start = nullptr;
    int32_t *d_s_cg_orig = nullptr, *d_s_cg_n = nullptr, *d_s_cg_scan = nullptr, *d_s_cg_start = nullptr;
    int32_t *d_s2o = nullptr, *d_s_atom_cg = nullptr, *d_s_type = nullptr;
    int64_t *d_vv_cnt = nullptr, *d_gen_cnt = nullptr, *d_vv_off = nullptr, *d_gen_off = nullptr;
    uint32_t *d_vv_ent = nullptr, *d_gen_ent = nullptr;
    int64_t vv_cap = 0, gen_cap = 0;
    uint64_t* d_counters = nullptr;
    int64_t *d_vv_cap = nullptr, *d_gen_cap = nullptr;
    int32_t *d_prev_vv = nullptr, *d_prev_gen = nullptr, *d_overflow = nullptr;
    bool have_prev = false;
    int64_t prev_tot_vv = 0, prev_tot_gen = 0;
    int64_t rebuilds = 0, refills = 0;
    double4* d_xq = nullptr;
    float4* d_xf = nullptr;
    long long *d_fi_s = nullptr, *d_fi_long_s = nullptr;
    int32_t nrow_blocks = 0;
    double *d_f_s = nullptr, *d_f_long_s = nullptr, *d_f_x = nullptr, *d_f_p = nullptr;
    double *d_rows = nullptr, *d_records = nullptr, *d_pert_e = nullptr;
    int64_t rec_cap = 0;
    int32_t* d_settle_fail = nullptr;
    uint8_t* d_pflag = nullptr;
    int32_t *d_pa_iac = nullptr, *d_pb_iac = nullptr, *d_plist = nullptr, *d_papairs = nullptr, *d_o2s = nullptr;
    double *d_pa_q = nullptr, *d_pb_q = nullptr, *d_palj = nullptr, *d_pacrf = nullptr;
    int32_t n_plist = 0, n_papairs = 0;
    char* d_temp = nullptr;
    size_t temp_bytes = 0;
    int cell_bits = 1;

    // state
    int cur = 0;
    bool built = false;
    int32_t steps_since_update = 0;
    int32_t nsnb = 5;
    bool twin = false, long_computed = false, has_pert = false;
    int32_t use_settle = 0, do_virial = 0;
    int64_t pair_count = 0;
    uint64_t host_counters[4] = {0, 0, 0, 0};
    // timing of the dominant kernel
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;

    ~gromos_amd_system() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    int alloc(T*& p, size_t n) {
        void* q = nullptr;
        size_t b = std::max<size_t>(n, 1) * sizeof(T);
        if (hipMalloc(&q, b) != hipSuccess) return fail(GROMOS_AMD_ENOMEM, "hipMalloc of %zu bytes failed", b);
        allocs.push_back(q);
        bytes += (int64_t)b;
        p = (T*)q;
        return 0;
    }
    template <class T>
    int realloc_dev(T*& p, size_t n, int64_t& cap) {
        if ((int64_t)n <= cap) return 0;
        size_t ncap = (size_t)(n * 1.25) + 1024;
        if (p) {
            auto it = std::find(allocs.begin(), allocs.end(), (void*)p);
            if (it != allocs.end()) allocs.erase(it);
            HIPCHK(hipStreamSynchronize(stream));
            HIPCHK(hipFree(p));
            bytes -= cap * (int64_t)sizeof(T);
            p = nullptr;
        }
        if (int rc = alloc(p, ncap)) return rc;
        cap = (int64_t)ncap;
        return 0;
    }
    template <class T>
    int upload(T*& p, const std::vector<T>& h) {
        if (int rc = alloc(p, h.size())) return rc;
        if (!h.empty()) HIPCHK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }

    SysPtrs ptrs() const {
        SysPtrs p{};
        p.jobs = d_jobs;
        p.cg_start = d_cg_start;
        p.cg_n = d_cg_n;
        p.cg_job = d_cg_job;
        p.atom_job = d_atom_job;
        p.q = d_q;
        p.iac = d_iac;
        p.mass = d_mass;
        p.inv_mass = d_inv_mass;
        p.lj = d_lj;
        p.excl_off = d_excl_off;
        p.excl = d_excl;
        p.o14_off = d_o14_off;
        p.o14 = d_o14;
        p.o14s_off = d_o14s_off;
        p.o14s = d_o14s;
        p.mol_first = d_mol_first;
        p.x_cur = d_x[cur];
        p.x_new = d_x[1 - cur];
        p.v_cur = d_v[cur];
        p.v_new = d_v[1 - cur];
        p.f = d_f;
        p.ref = d_ref;
        p.key_in = d_key_in;
        p.val_in = d_val_in;
        p.key_out = d_key_out;
        p.val_out = d_val_out;
        p.cell_count = d_cell_count;
        p.cell_start = d_cell_start;
        p.s_cg_orig = d_s_cg_orig;
        p.s_cg_n = d_s_cg_n;
        p.s_cg_scan = d_s_cg_scan;
        p.s_cg_start = d_s_cg_start;
        p.s_ref = d_s_ref;
        p.s2o = d_s2o;
        p.s_atom_cg = d_s_atom_cg;
        p.s_type = d_s_type;
        p.vv_cnt = d_vv_cnt;
        p.gen_cnt = d_gen_cnt;
        p.vv_off = d_vv_off;
        p.gen_off = d_gen_off;
        p.vv_ent = d_vv_ent;
        p.gen_ent = d_gen_ent;
        p.counters = d_counters;
        p.vv_cap = d_vv_cap;
        p.gen_cap = d_gen_cap;
        p.prev_vv = d_prev_vv;
        p.prev_gen = d_prev_gen;
        p.overflow = d_overflow;
        p.xq = d_xq;
        p.xf = d_xf;
        p.fi_s = d_fi_s;
        p.fi_long_s = d_fi_long_s;
        p.nrow_blocks = nrow_blocks;
        p.f_s = d_f_s;
        p.f_long_s = d_f_long_s;
        p.f_x = d_f_x;
        p.f_p = d_f_p;
        p.rows = d_rows;
        p.records = d_records;
        p.pert_e = d_pert_e;
        p.settle_fail = d_settle_fail;
        p.pflag = d_pflag;
        p.pa_iac = d_pa_iac;
        p.pb_iac = d_pb_iac;
        p.pa_q = d_pa_q;
        p.pb_q = d_pb_q;
        p.palj = d_palj;
        p.pacrf = d_pacrf;
        p.plist = d_plist;
        p.n_plist = n_plist;
        p.papairs = d_papairs;
        p.n_papairs = n_papairs;
        p.o2s = d_o2s;
        return p;
    }

    hipEvent_t next_event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }

    int collect_timing() {
        if (ev_used == 0) return 0;
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t k = 0; k + 1 < ev_used; k += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_pool[k], ev_pool[k + 1]));
            timed_ms += ms;
            timed_launches++;
        }
        ev_used = 0;
        return 0;
    }
};

static int build_pairlist(gromos_amd_system* S);

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// cells per axis: the largest count whose edge still exceeds the list cutoff
// (with a relative margin so that rounding in the wrapped coordinates can never
// push an in-range pair two cells apart); 1 when the box is smaller.
static int32_t cells_for(double L, double rl) {
    if (!(rl > 0.0)) return 1;
    int64_t n = (int64_t)std::floor(L / (rl * (1.0 + 1e-9) + 1e-12));
    if (n < 1) n = 1;
    if (n > 1024) n = 1024;
    return (int32_t)n;
}

static int create_batch(const gromos_amd_topology* topos, const gromos_amd_params* prms, int njobs,
                        int32_t precision, gromos_amd_system** out) {
    if (!out || njobs <= 0 || !topos || !prms) return fail(GROMOS_AMD_EINVAL, "create: bad arguments");
    *out = nullptr;
    auto* S = new gromos_amd_system();
    std::unique_ptr<gromos_amd_system> guard(S);
    S->precision = precision;
    S->njobs = njobs;
    int64_t atom_off = 0, cg_off = 0, lj_off = 0, cell_off = 0;
    std::vector<int32_t> cg_start, cg_n, cg_job, atom_job, iac, excl_off, excl, o14_off, o14, o14s_off, o14s,
        mol_first;
    std::vector<double> q, mass, inv_mass, lj;
    std::vector<uint8_t> pflag;
    std::vector<int32_t> pa_iac, pb_iac, plist, papairs;
    std::vector<double> pa_q, pb_q, palj, pacrf;
    excl_off.push_back(0);
    o14_off.push_back(0);
    o14s_off.push_back(0);
    const gromos_amd_params& P0 = prms[0];
    for (int j = 0; j < njobs; ++j) {
        const gromos_amd_topology& T = topos[j];
        const gromos_amd_params& P = prms[j];
        if (int rc = check_box(P.boundary, P.box)) return rc;
        int32_t n = T.n_atoms;
        if (n <= 0 || T.n_cg <= 0 || !T.cg_start || !T.charge || !T.iac || !T.mass || !T.lj || T.n_types <= 0 ||
            !T.excl_off || !T.o14_off)
            return fail(GROMOS_AMD_EINVAL, "job %d: incomplete topology", j);
        if (T.cg_start[0] != 0 || T.cg_start[T.n_cg] != n) return fail(GROMOS_AMD_EINVAL, "job %d: cg_start must span [0, n_atoms]", j);
        for (int c = 0; c < T.n_cg; ++c)
            if (T.cg_start[c + 1] <= T.cg_start[c]) return fail(GROMOS_AMD_EINVAL, "job %d: empty charge group %d", j, c);
        if (T.n_solute_cg < 0 || T.n_solute_cg > T.n_cg || T.n_solute_atoms != T.cg_start[T.n_solute_cg])
            return fail(GROMOS_AMD_EINVAL, "job %d: solute charge groups must end at n_solute_atoms", j);
        for (int a = 0; a < n; ++a)
            if (T.iac[a] < 0 || T.iac[a] >= T.n_types) return fail(GROMOS_AMD_EINVAL, "job %d: atom %d type out of range", j, a);
        int32_t nsolv_atoms = n - T.n_solute_atoms;
        int32_t aps = T.atoms_per_solvent;
        if (nsolv_atoms > 0) {
            if (aps < 1 || aps > 4) return fail(GROMOS_AMD_EINVAL, "job %d: atoms_per_solvent must be 1..4", j);
            for (int c = T.n_solute_cg; c < T.n_cg; ++c)
                if (T.cg_start[c + 1] - T.cg_start[c] != aps) return fail(GROMOS_AMD_EINVAL, "job %d: solvent group size != atoms_per_solvent", j);
        }
        if (P.nsnb != P0.nsnb) return fail(GROMOS_AMD_EINVAL, "jobs of one batch share the rebuild cadence");
        if (nsolv_atoms > 0 && j > 0 && topos[0].n_atoms > topos[0].n_solute_atoms && aps != topos[0].atoms_per_solvent)
            return fail(GROMOS_AMD_EINVAL, "jobs of one batch share the solvent molecule size");
        if (P.rcutl < P.rcutp || !(P.rcutl > 0.0)) return fail(GROMOS_AMD_EINVAL, "job %d: need 0 < rcutp <= rcutl", j);
        JobDev J{};
        J.box = make_box(P.boundary, P.box);
        gromos_amd_crf c;
        gromos_amd_crf_new(P.rcutl, 1.0, P.epsrf, P.kappa, &c);   // md.rs:326-329, 815
        J.crf = to_crf(&c);
        J.fpepsi = P.four_pi_eps_i;
        double s = P.rcutp + P.skin, l = P.rcutl + P.skin;
        J.c2s = s * s;   // (rc + skin).powi(2), pairlist.rs:192-193
        J.c2l = l * l;
        J.n_types = T.n_types;
        J.lj_off = (int32_t)lj_off;
        J.aps = aps;
        J.atom_off = (int32_t)atom_off;
        J.n_atoms = n;
        J.n_solute_atoms = T.n_solute_atoms;
        J.cg_off = (int32_t)cg_off;
        J.n_cg = T.n_cg;
        J.n_solute_cg = T.n_solute_cg;
        J.twin = P.rcutp < P.rcutl - 1e-10;
        J.dt = P.dt;
        J.dist_oh = T.dist_oh;
        J.dist_hh = T.dist_hh;
        // use_cg_grouped_solute (forcefield.rs:271-283)
        if (J.box.kind == 0) J.use_cg_grouped = 1;
        else {
            double hm = std::min(std::min(0.5 * P.box[0], 0.5 * P.box[1]), 0.5 * P.box[2]);
            J.use_cg_grouped = hm > P.rcutl + P.skin + 0.3;
        }
        for (int a = 0; a < 3; ++a) J.ncell[a] = J.box.kind ? cells_for(P.box[a], l) : 1;
        J.cell_off = (int32_t)cell_off;
        cell_off += (int64_t)J.ncell[0] * J.ncell[1] * J.ncell[2];
        // solvent template tables from the first solvent molecule (innerloops.rs:241-255)
        if (nsolv_atoms > 0) {
            int32_t f = T.n_solute_atoms;
            for (int ai = 0; ai < aps; ++ai)
                for (int aj = 0; aj < aps; ++aj) {
                    const double* p = T.lj + 4 * ((int64_t)T.iac[f + ai] * T.n_types + T.iac[f + aj]);
                    J.sv_c6[ai * 4 + aj] = p[0];
                    J.sv_c12[ai * 4 + aj] = p[1];
                    J.sv_qq[ai * 4 + aj] = T.charge[f + ai] * T.charge[f + aj] * P.four_pi_eps_i;
                }
            J.m_o = T.mass[f];
            J.m_h = aps > 1 ? T.mass[f + 1] : 0.0;
        }
        if (T.pert_flag) {
            for (int a = 0; a < n; ++a)
                if (T.pert_flag[a]) {
                    J.has_pert = 1;
                    if (a >= T.n_solute_atoms) return fail(GROMOS_AMD_EINVAL, "job %d: only solute atoms can be perturbed", j);
                    if (!T.pert_a_iac || !T.pert_b_iac || !T.pert_a_q || !T.pert_b_q || !T.pert_alpha_lj || !T.pert_alpha_crf)
                        return fail(GROMOS_AMD_EINVAL, "job %d: incomplete perturbation arrays", j);
                    if (T.pert_a_iac[a] < 0 || T.pert_a_iac[a] >= T.n_types || T.pert_b_iac[a] < 0 || T.pert_b_iac[a] >= T.n_types)
                        return fail(GROMOS_AMD_EINVAL, "job %d: perturbed type of atom %d out of range", j, a);
                }
        }
        // from_lambda(l, l, l, l, 1, 1, 1, 1, NLAM) (forcefield.rs:767-777)
        {
            gromos_amd_lambda L;
            gromos_amd_lambda_from(P.lambda, P.lambda, P.lambda, P.lambda, 1.0, 1.0, 1.0, 1.0, P.nlam, &L);
            std::memcpy(&J.lp, &L, sizeof(L));
        }
        if (P.use_settle && nsolv_atoms > 0) {
            if (aps != 3) return fail(GROMOS_AMD_EINVAL, "SETTLE needs 3-atom solvent (settle_algorithm.rs:43-45)");
            if (T.mass[T.n_solute_atoms + 1] != T.mass[T.n_solute_atoms + 2])
                return fail(GROMOS_AMD_EINVAL, "SETTLE needs equal H masses (settle_algorithm.rs:46-51)");
            for (int32_t a = T.n_solute_atoms; a + 3 <= n; a += 3) mol_first.push_back((int32_t)atom_off + a);
        }
        S->jobs.push_back(J);
        S->job_n_atoms.push_back(n);
        // arrays
        for (int c = 0; c < T.n_cg; ++c) {
            cg_start.push_back((int32_t)atom_off + T.cg_start[c]);
            cg_n.push_back(T.cg_start[c + 1] - T.cg_start[c]);
            cg_job.push_back(j);
        }
        int64_t npad_j = align_up(n, BLOCK);
        std::vector<std::vector<int32_t>> sym(n);
        for (int a = 0; a < npad_j; ++a) {
            bool live = a < n;
            atom_job.push_back(live ? j : -1);
            q.push_back(live ? T.charge[a] : 0.0);
            iac.push_back(live ? T.iac[a] : 0);
            mass.push_back(live ? T.mass[a] : 0.0);
            inv_mass.push_back(live && T.mass[a] > 0.0 ? 1.0 / T.mass[a] : 0.0);
            if (live) {
                for (int k = T.excl_off[a]; k < T.excl_off[a + 1]; ++k) {
                    int32_t b = T.excl[k];
                    if (b < 0 || b >= n) return fail(GROMOS_AMD_EINVAL, "job %d: exclusion index out of range", j);
                    excl.push_back((int32_t)atom_off + b);
                }
                for (int k = T.o14_off[a]; k < T.o14_off[a + 1]; ++k) {
                    int32_t b = T.o14[k];
                    if (b < 0 || b >= n) return fail(GROMOS_AMD_EINVAL, "job %d: 1-4 index out of range", j);
                    o14.push_back((int32_t)atom_off + b);
                    if (b > a) {  // excluded.rs:148-151 evaluates j > i only
                        sym[a].push_back((int32_t)(atom_off + b) | (int32_t)0x80000000u);
                        sym[b].push_back((int32_t)(atom_off + a));
                    }
                }
            }
            bool pa = live && T.pert_flag && T.pert_flag[a];
            pflag.push_back(pa ? 1 : 0);
            pa_iac.push_back(pa ? T.pert_a_iac[a] : 0);
            pb_iac.push_back(pa ? T.pert_b_iac[a] : 0);
            pa_q.push_back(pa ? T.pert_a_q[a] : 0.0);
            pb_q.push_back(pa ? T.pert_b_q[a] : 0.0);
            palj.push_back(pa ? T.pert_alpha_lj[a] : 0.0);
            pacrf.push_back(pa ? T.pert_alpha_crf[a] : 0.0);
            if (pa) plist.push_back((int32_t)atom_off + a);
            excl_off.push_back((int32_t)excl.size());
            o14_off.push_back((int32_t)o14.size());
            if (live) for (int32_t e : sym[a]) o14s.push_back(e);
            o14s_off.push_back((int32_t)o14s.size());
        }
        for (int64_t k = 0; k < (int64_t)T.n_types * T.n_types * 4; ++k) lj.push_back(T.lj[k]);
        if (T.pert_flag && T.n_pert_atom_pairs > 0) {
            for (int k = 0; k < T.n_pert_atom_pairs; ++k) {
                const int32_t* e = T.pert_atom_pairs + 4 * k;
                if (e[0] < 0 || e[0] >= n || e[1] < 0 || e[1] >= n) return fail(GROMOS_AMD_EINVAL, "job %d: atom pair out of range", j);
                papairs.push_back((int32_t)atom_off + e[0]);
                papairs.push_back((int32_t)atom_off + e[1]);
                papairs.push_back(e[2]);
                papairs.push_back(e[3]);
            }
        }
        lj_off += (int64_t)T.n_types * T.n_types * 4;
        atom_off += npad_j;
        cg_off += T.n_cg;
    }
    S->npad = (int32_t)atom_off;
    S->ncg = (int32_t)cg_off;
    S->ncell = (int32_t)cell_off;
    S->nmol = (int32_t)mol_first.size();
    S->nrow_blocks = (int32_t)(align_up((int64_t)S->npad * 4, BLOCK) / BLOCK);   // 4 lanes per atom in the pair kernel
    if (S->npad >= (int32_t)ENT_IDX) return fail(GROMOS_AMD_EINVAL, "too many atoms");
    S->nsnb = std::max(1, P0.nsnb);
    S->use_settle = P0.use_settle;
    S->do_virial = P0.do_virial;
    for (auto& J : S->jobs) S->twin = S->twin || J.twin;
    while ((1ll << S->cell_bits) < (int64_t)S->ncell + 1) S->cell_bits++;
    if (S->ncg >= (int32_t)ENT_IDX) return fail(GROMOS_AMD_EINVAL, "too many charge groups");

    HIPCHK(hipStreamCreate(&S->stream));
    int rc = 0;
    std::vector<JobDev> jobs = S->jobs;
    rc = rc ? rc : S->upload(S->d_jobs, jobs);
    rc = rc ? rc : S->upload(S->d_cg_start, cg_start);
    rc = rc ? rc : S->upload(S->d_cg_n, cg_n);
    rc = rc ? rc : S->upload(S->d_cg_job, cg_job);
    rc = rc ? rc : S->upload(S->d_atom_job, atom_job);
    rc = rc ? rc : S->upload(S->d_q, q);
    rc = rc ? rc : S->upload(S->d_iac, iac);
    rc = rc ? rc : S->upload(S->d_mass, mass);
    rc = rc ? rc : S->upload(S->d_inv_mass, inv_mass);
    rc = rc ? rc : S->upload(S->d_lj, lj);
    rc = rc ? rc : S->upload(S->d_excl_off, excl_off);
    rc = rc ? rc : S->upload(S->d_excl, excl);
    rc = rc ? rc : S->upload(S->d_o14_off, o14_off);
    rc = rc ? rc : S->upload(S->d_o14, o14);
    rc = rc ? rc : S->upload(S->d_o14s_off, o14s_off);
    rc = rc ? rc : S->upload(S->d_o14s, o14s);
    rc = rc ? rc : S->upload(S->d_mol_first, mol_first);
    rc = rc ? rc : S->upload(S->d_pflag, pflag);
    rc = rc ? rc : S->upload(S->d_pa_iac, pa_iac);
    rc = rc ? rc : S->upload(S->d_pb_iac, pb_iac);
    rc = rc ? rc : S->upload(S->d_pa_q, pa_q);
    rc = rc ? rc : S->upload(S->d_pb_q, pb_q);
    rc = rc ? rc : S->upload(S->d_palj, palj);
    rc = rc ? rc : S->upload(S->d_pacrf, pacrf);
    rc = rc ? rc : S->upload(S->d_plist, plist);
    rc = rc ? rc : S->upload(S->d_papairs, papairs);
    S->n_plist = (int32_t)plist.size();
    S->n_papairs = (int32_t)papairs.size() / 4;
    S->has_pert = S->n_plist > 0;
    if (rc) return rc;
    size_t n3 = (size_t)S->npad * 3, ncg = (size_t)S->ncg;
    for (int k = 0; k < 2; ++k) {
        if ((rc = S->alloc(S->d_x[k], n3))) return rc;
        if ((rc = S->alloc(S->d_v[k], n3))) return rc;
        HIPCHK(hipMemset(S->d_x[k], 0, n3 * sizeof(double)));
        HIPCHK(hipMemset(S->d_v[k], 0, n3 * sizeof(double)));
    }
    if ((rc = S->alloc(S->d_f, n3)) || (rc = S->alloc(S->d_ref, 3 * ncg)) || (rc = S->alloc(S->d_s_ref, 3 * ncg)) ||
        (rc = S->alloc(S->d_key_in, ncg)) || (rc = S->alloc(S->d_val_in, ncg)) || (rc = S->alloc(S->d_key_out, ncg)) ||
        (rc = S->alloc(S->d_val_out, ncg)) || (rc = S->alloc(S->d_cell_count, (size_t)S->ncell + 1)) ||
        (rc = S->alloc(S->d_cell_start, (size_t)S->ncell + 1)) || (rc = S->alloc(S->d_s_cg_orig, ncg)) ||
        (rc = S->alloc(S->d_s_cg_n, ncg + 1)) || (rc = S->alloc(S->d_s_cg_scan, ncg + 1)) ||
        (rc = S->alloc(S->d_s_cg_start, ncg)) || (rc = S->alloc(S->d_s2o, (size_t)S->npad)) ||
        (rc = S->alloc(S->d_s_atom_cg, (size_t)S->npad)) || (rc = S->alloc(S->d_o2s, (size_t)S->npad)) || (rc = S->alloc(S->d_s_type, (size_t)S->npad)) ||
        (rc = S->alloc(S->d_vv_cnt, ncg + 1)) || (rc = S->alloc(S->d_gen_cnt, ncg + 1)) ||
        (rc = S->alloc(S->d_vv_off, ncg + 1)) || (rc = S->alloc(S->d_gen_off, ncg + 1)) ||
        (rc = S->alloc(S->d_counters, 4)) || (rc = S->alloc(S->d_xq, (size_t)S->npad)) ||
        (rc = S->alloc(S->d_vv_cap, ncg + 1)) || (rc = S->alloc(S->d_gen_cap, ncg + 1)) ||
        (rc = S->alloc(S->d_prev_vv, ncg)) || (rc = S->alloc(S->d_prev_gen, ncg)) ||
        (rc = S->alloc(S->d_overflow, 1)) ||
        (rc = S->alloc(S->d_f_s, n3)) || (rc = S->alloc(S->d_f_long_s, n3)) || (rc = S->alloc(S->d_f_x, n3)) ||
        (rc = S->alloc(S->d_f_p, n3)) || (rc = S->alloc(S->d_xf, (size_t)S->npad)) ||
        (rc = S->alloc(S->d_fi_s, n3)) || (rc = S->alloc(S->d_fi_long_s, n3)) ||
        (rc = S->alloc(S->d_rows, (size_t)NSLOT * S->nrow_blocks * ROW)) ||
        (rc = S->alloc(S->d_pert_e, (size_t)3 * njobs)) || (rc = S->alloc(S->d_settle_fail, (size_t)njobs)))
        return rc;
    HIPCHK(hipMemset(S->d_f_long_s, 0, n3 * sizeof(double)));
    HIPCHK(hipMemset(S->d_f_p, 0, n3 * sizeof(double)));
    HIPCHK(hipMemset(S->d_pert_e, 0, 3 * njobs * sizeof(double)));
    HIPCHK(hipMemset(S->d_rows, 0, (size_t)NSLOT * S->nrow_blocks * ROW * sizeof(double)));
    HIPCHK(hipMemset(S->d_fi_long_s, 0, n3 * sizeof(long long)));
    HIPCHK(hipMemset(S->d_s_cg_n, 0, (ncg + 1) * sizeof(int32_t)));
    HIPCHK(hipMemset(S->d_vv_cnt, 0, (ncg + 1) * sizeof(int64_t)));
    HIPCHK(hipMemset(S->d_gen_cnt, 0, (ncg + 1) * sizeof(int64_t)));
    HIPCHK(hipMemset(S->d_vv_cap, 0, (ncg + 1) * sizeof(int64_t)));
    HIPCHK(hipMemset(S->d_gen_cap, 0, (ncg + 1) * sizeof(int64_t)));
    // scratch for the library sort / scans
    size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, S->d_key_in, S->d_key_out, S->d_val_in, S->d_val_out,
                                              S->ncg, 0, S->cell_bits, S->stream));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, S->d_cell_count, S->d_cell_start, S->ncell + 1, S->stream));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, S->d_s_cg_n, S->d_s_cg_scan, S->ncg + 1, S->stream));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, S->d_vv_cnt, S->d_vv_off, S->ncg + 1, S->stream));
    S->temp_bytes = std::max(std::max(t1, t2), std::max(t3, t4));
    if ((rc = S->alloc(S->d_temp, S->temp_bytes))) return rc;
    guard.release();
    *out = S;
    return 0;
}

extern "C" int gromos_amd_system_create(const gromos_amd_topology* topo, const gromos_amd_params* prm,
                                        int32_t precision, gromos_amd_system** out) {
    if (precision != GROMOS_AMD_FP64 && precision != GROMOS_AMD_MIXED)
        return fail(GROMOS_AMD_EINVAL, "precision must be 0 (FP64) or 1 (mixed)");
    return create_batch(topo, prm, 1, precision, out);
}

extern "C" void gromos_amd_system_destroy(gromos_amd_system* S) { delete S; }

extern "C" int64_t gromos_amd_system_device_bytes(const gromos_amd_system* S) { return S ? S->bytes : 0; }
extern "C" int64_t gromos_amd_system_pair_count(const gromos_amd_system* S) { return S ? S->pair_count : 0; }

extern "C" int gromos_amd_system_set_state(gromos_amd_system* S, const double* pos, const double* vel) {
    if (!S || !pos) return fail(GROMOS_AMD_EINVAL, "set_state: null argument");
    for (int j = 0; j < S->njobs; ++j) {
        size_t off = (size_t)S->jobs[j].atom_off * 3, n3 = (size_t)S->job_n_atoms[j] * 3;
        size_t hoff = 0;
        for (int k = 0; k < j; ++k) hoff += (size_t)S->job_n_atoms[k] * 3;
        HIPCHK(hipMemcpyAsync(S->d_x[S->cur] + off, pos + hoff, n3 * sizeof(double), hipMemcpyHostToDevice, S->stream));
        if (vel) HIPCHK(hipMemcpyAsync(S->d_v[S->cur] + off, vel + hoff, n3 * sizeof(double), hipMemcpyHostToDevice, S->stream));
        else HIPCHK(hipMemsetAsync(S->d_v[S->cur] + off, 0, n3 * sizeof(double), S->stream));
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    S->built = false;
    S->long_computed = false;
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

// StandardPairlistAlgorithm::update (pairlist.rs:155-174) on the device.
// Cell sort, then the neighbour lists.  After the first build the lists are
// filled in one pass into slots sized from each group's previous count; if a
// group outgrew its slot (flag) the lists are refilled with the exact counts
// that pass produced.
static int scan_offsets(gromos_amd_system* S, const int64_t* cnt_vv, const int64_t* cnt_gen) {
    size_t tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, cnt_vv, S->d_vv_off, S->ncg + 1, S->stream));
    tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, cnt_gen, S->d_gen_off, S->ncg + 1, S->stream));
    return 0;
}

static int build_pairlist(gromos_amd_system* S) {
    hipStream_t st = S->stream;
    SysPtrs p = S->ptrs();
    int32_t ncg = S->ncg;
    launch_cg_ref(p, ncg, st);
    HIPCHK(hipMemsetAsync(S->d_cell_count, 0, ((size_t)S->ncell + 1) * sizeof(int32_t), st));
    launch_cell_hist(p, ncg, st);
    size_t tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, S->d_cell_count, S->d_cell_start, S->ncell + 1, st));
    tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(S->d_temp, tb, S->d_key_in, S->d_key_out, S->d_val_in, S->d_val_out,
                                              ncg, 0, S->cell_bits, st));
    launch_sorted_cg(p, ncg, st);
    tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, S->d_s_cg_n, S->d_s_cg_scan, ncg + 1, st));
    HIPCHK(hipMemsetAsync(S->d_s2o, 0xFF, (size_t)S->npad * sizeof(int32_t), st));
    launch_sorted_atoms(p, ncg, st);
    HIPCHK(hipMemsetAsync(S->d_counters, 0, 4 * sizeof(uint64_t), st));
    HIPCHK(hipMemsetAsync(S->d_overflow, 0, sizeof(int32_t), st));
    int32_t overflow = 0;
    bool need_exact = !S->have_prev;
    if (S->have_prev) {
        launch_caps(p, ncg, 0, st);
        if (int rc = scan_offsets(S, S->d_vv_cap, S->d_gen_cap)) return rc;
        int64_t bv = S->prev_tot_vv + S->prev_tot_vv / 8 + 16 * (int64_t)ncg + 64;
        int64_t bg = S->prev_tot_gen + S->prev_tot_gen / 8 + 16 * (int64_t)ncg + 64;
        if (int rc = S->realloc_dev(S->d_vv_ent, (size_t)bv, S->vv_cap)) return rc;
        if (int rc = S->realloc_dev(S->d_gen_ent, (size_t)bg, S->gen_cap)) return rc;
        p = S->ptrs();
        launch_nbl(p, ncg, true, st);
        HIPCHK(hipMemcpyAsync(&overflow, S->d_overflow, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(S->host_counters, S->d_counters, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (overflow) {
            S->refills++;
            need_exact = true;   // the fill pass left exact counts in vv_cnt / gen_cnt
        }
    } else {
        launch_nbl(p, ncg, false, st);
    }
    if (need_exact) {
        if (int rc = scan_offsets(S, S->d_vv_cnt, S->d_gen_cnt)) return rc;
        int64_t tot[2];
        HIPCHK(hipMemcpyAsync(&tot[0], S->d_vv_off + ncg, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&tot[1], S->d_gen_off + ncg, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(S->host_counters, S->d_counters, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (int rc = S->realloc_dev(S->d_vv_ent, (size_t)tot[0], S->vv_cap)) return rc;
        if (int rc = S->realloc_dev(S->d_gen_ent, (size_t)tot[1], S->gen_cap)) return rc;
        p = S->ptrs();
        launch_caps(p, ncg, 1, st);
        uint64_t keep[4];
        std::memcpy(keep, S->host_counters, sizeof(keep));
        launch_nbl(p, ncg, true, st);
        std::memcpy(S->host_counters, keep, sizeof(keep));
    }
    launch_caps(p, ncg, 2, st);
    HIPCHK(hipGetLastError());
    S->prev_tot_vv = (int64_t)S->host_counters[2];
    S->prev_tot_gen = (int64_t)S->host_counters[3];
    S->have_prev = true;
    S->built = true;
    S->rebuilds++;
    S->steps_since_update = 0;
    return 0;
}

extern "C" int gromos_amd_system_pairlist_update(gromos_amd_system* S) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    return build_pairlist(S);
}

// Forcefield::apply nonbonded part (forcefield.rs:226-1060) into d_f, energies into
// records[step]
static int forcefield(gromos_amd_system* S, int32_t rec_step, bool with_kin_rows) {
    hipStream_t st = S->stream;
    if (!S->built) {
        if (int rc = build_pairlist(S)) return rc;
    }
    bool rebuild = S->steps_since_update >= S->nsnb;   // pairlist.rs:105-107
    if (rebuild) {
        if (int rc = build_pairlist(S)) return rc;
    }
    S->steps_since_update++;
    SysPtrs p = S->ptrs();
    int32_t npad = S->npad;
    const int32_t mixed = S->precision == GROMOS_AMD_MIXED;
    launch_gather(p, npad, mixed, st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (S->timing) {
        e0 = S->next_event();
        e1 = S->next_event();
        HIPCHK(hipEventRecord(e0, st));
    }
    int32_t aps = S->jobs[0].aps > 0 ? S->jobs[0].aps : 3;
    launch_vv(p, npad, aps, S->do_virial, mixed, st);
    if (S->timing) HIPCHK(hipEventRecord(e1, st));
    launch_gen(p, npad, false, S->do_virial, mixed, st);
    S->pair_count += (int64_t)S->host_counters[0];
    if (S->twin && (rebuild || !S->long_computed)) {
        launch_gen(p, npad, true, 1, mixed, st);
        S->long_computed = true;
        S->pair_count += (int64_t)S->host_counters[1];
    }
    launch_excl14(p, npad, st);
    if (S->has_pert) launch_pert(p, st);
    launch_assemble(p, npad, S->twin ? 1 : 0, S->has_pert ? 1 : 0, mixed, st);
    (void)with_kin_rows;
    (void)rec_step;
    HIPCHK(hipGetLastError());
    return 0;
}

static int ensure_records(gromos_amd_system* S, int64_t nsteps) {
    int64_t need = std::max<int64_t>(nsteps, 1) * S->njobs * REC;
    return S->realloc_dev(S->d_records, (size_t)need, S->rec_cap);
}

extern "C" int gromos_amd_system_forcefield_apply(gromos_amd_system* S, double* forces, double* energies,
                                                  double* virial) {
    if (!S) return fail(GROMOS_AMD_EINVAL, "null system");
    if (int rc = ensure_records(S, 1)) return rc;
    if (int rc = forcefield(S, 0, false)) return rc;
    SysPtrs p = S->ptrs();
    launch_reduce(p, S->npad, S->twin ? 1 : 0, 0, 0, S->njobs, S->stream);
    HIPCHK(hipGetLastError());
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

extern "C" int gromos_amd_system_md(gromos_amd_system* S, int32_t nsteps, double* energies) {
    if (!S || nsteps < 0) return fail(GROMOS_AMD_EINVAL, "md: bad arguments");
    if (nsteps == 0) return 0;
    if (int rc = ensure_records(S, nsteps)) return rc;
    std::vector<int32_t> big(S->njobs, 0x7FFFFFFF);
    HIPCHK(hipMemcpyAsync(S->d_settle_fail, big.data(), big.size() * sizeof(int32_t), hipMemcpyHostToDevice, S->stream));
    for (int32_t s = 0; s < nsteps; ++s) {
        if (int rc = forcefield(S, s, true)) return rc;
        SysPtrs p = S->ptrs();
        launch_leapfrog(p, S->npad, S->stream);
        if (S->use_settle) launch_settle(p, S->nmol, s, S->stream);
        launch_ekin(p, S->npad, S->stream);
        launch_reduce(p, S->npad, S->twin ? 1 : 0, 1, s, S->njobs, S->stream);
        S->cur = 1 - S->cur;   // exchange_state (configuration.rs:426-428)
    }
    HIPCHK(hipGetLastError());
    std::vector<int32_t> failed(S->njobs);
    HIPCHK(hipMemcpyAsync(failed.data(), S->d_settle_fail, failed.size() * sizeof(int32_t), hipMemcpyDeviceToHost, S->stream));
    std::vector<double> rec;
    if (energies) {
        rec.resize((size_t)nsteps * S->njobs * REC);
        HIPCHK(hipMemcpyAsync(rec.data(), S->d_records, rec.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    }
    HIPCHK(hipStreamSynchronize(S->stream));
    if (energies) {
        for (int32_t s = 0; s < nsteps; ++s)
            for (int j = 0; j < S->njobs; ++j) {
                const double* r = rec.data() + ((size_t)s * S->njobs + j) * REC;
                double* e = energies + ((size_t)s * S->njobs + j) * 5;
                for (int c = 0; c < 5; ++c) e[c] = r[c];
            }
    }
    if (int rc = S->collect_timing()) return rc;
    for (int j = 0; j < S->njobs; ++j)
        if (failed[j] != 0x7FFFFFFF)
            return fail(GROMOS_AMD_ESETTLE, "SETTLE failed: sin(angle) > 1.0 in analytical solve (job %d, step %d)", j, failed[j]);
    return 0;
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
    // gather positions so that the materializer sees the current sorted layout
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
        }
        S->timing = on;
    }
    return 0;
}
