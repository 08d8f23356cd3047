// Host side of the device-resident systems: error reporting, the per-job host
// image of a topology, the batch (gromos_amd_system) with its device buffers,
// and the step driver shared by the single-system C ABI and the session.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gromos_amd.h"
#include "bonded.hpp"
#include "system_kernels.hpp"

namespace gamd {

// last_error() text of the calling thread; every entry point returns the code
int fail(int code, const char* fmt, ...);

#define HIPCHK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return ::gamd::fail(GROMOS_AMD_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

Box make_box(int32_t boundary, const double* box);
int check_box(int32_t boundary, const double* box);
inline Crf to_crf(const gromos_amd_crf* c) { return Crf{c->crf_cut, c->crf_2cut3i, c->crf_cut3i, c->cutoff_sq}; }
inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct BondedTerms {
    std::vector<int4> atoms;
    std::vector<double4> prm;
    std::vector<int32_t> kind;
    int32_t n[4] = {0, 0, 0, 0};
};

// One job's topology and run parameters, validated and flattened once (local
// atom / group indices); a batch layout concatenates images with offsets.
struct JobImage {
    JobDev J{};                     // constants; offsets are set at layout
    gromos_amd_params P{};
    int32_t n = 0, n_cg = 0, n_types = 0;
    bool solvent = false;           // has solvent molecules
    std::vector<int32_t> cg_start;  // n_cg + 1
    std::vector<double> q, mass, lj;
    std::vector<int32_t> iac;
    std::vector<int32_t> excl_off, excl;
    std::vector<int32_t> o14_off, o14;
    std::vector<int32_t> o14s_off, o14s;   // symmetric 1-4, bit 31 = owner is the lower atom
    std::vector<int32_t> mol_first;        // SETTLE waters
    bool has_pert = false;
    std::vector<uint8_t> pflag;
    std::vector<int32_t> pa_iac, pb_iac, papairs;
    std::vector<double> pa_q, pb_q, palj, pacrf;
    BondedTerms bonded;             // covalent terms of the solute
    std::vector<int32_t> pg_start;  // pressure groups (job atoms), n_pg + 1; empty: none
    struct Constraint {
        int32_t i, j;
        double len;
    };
    std::vector<Constraint> shake;  // the solute's SHAKE constraints (job atoms)
    int64_t footprint = 0;          // device bytes this job adds to a batch (estimate)
    int64_t list_entries = 0;       // full neighbour-list entries of the job (estimate)
};

// The pair kernels index the neighbour lists with 32-bit offsets: a batch's
// list entries (either list) must stay below this.
constexpr int64_t MAX_LIST_ENTRIES = INT32_MAX;

int make_image(const gromos_amd_topology& T, const gromos_amd_params& P, int job, JobImage& out);
// a pressure-coupled job's image takes the box it has now: edge lengths, the
// cell grid (5 % to spare) and the CG-grouped predicate follow
void rebox_image(JobImage& im, const Box& box);

// validated covalent terms of one topology, kind-major (bonds, angles,
// dihedrals, impropers), local atom indices; n: terms kept per kind
int bonded_terms(const gromos_amd_bonded& b, int32_t n_atoms, int job, BondedTerms& out);
// per-atom CSR of the terms' contributions (entries term << 2 | slot, term order)
void bonded_csr(const BondedTerms& t, int32_t natoms, std::vector<int32_t>& aoff, std::vector<int32_t>& aref);
// GROMOS_AMD_ENOMEM unless both list sizes fit the kernels' 32-bit offsets
int check_list_size(int64_t vv_entries, int64_t gen_entries);

}  // namespace gamd

// Pinned host staging for a batch layout's arrays (hipHostMalloc: DMA uploads and
// no page faults when reused); carved afresh by every layout, grown between layouts.
struct PinnedArena {
    template <class T>
    struct Span {
        T* p = nullptr;
        size_t n = 0;
        T& operator[](size_t i) const { return p[i]; }
        size_t size() const { return n; }
        T* data() const { return p; }
        T* begin() const { return p; }
        T* end() const { return p + n; }
    };
    char* base = nullptr;
    size_t cap = 0, used = 0;
    ~PinnedArena() {
        if (base) (void)hipHostFree(base);
    }
    PinnedArena() = default;
    PinnedArena(const PinnedArena&) = delete;
    PinnedArena& operator=(const PinnedArena&) = delete;
    // room for `bytes` in `parts` arrays; the caller has synchronised with every
    // copy still reading the old contents
    int reset(size_t bytes, size_t parts) {
        const size_t need = bytes + (parts + 1) * 256;
        if (need > cap) {
            // an arena that grows at least doubles (pinning is slow; a session's
            // batches grow job by job)
            size_t c = cap ? std::max(2 * need, 2 * cap) : need + need / 4;
            if (base) HIPCHK(hipHostFree(base));
            base = nullptr;
            cap = 0;
            if (hipHostMalloc((void**)&base, c, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                base = nullptr;
                c = need + need / 4;   // the plain slack before giving up
                if (hipHostMalloc((void**)&base, c, hipHostMallocDefault) != hipSuccess) {
                    base = nullptr;
                    return ::gamd::fail(GROMOS_AMD_ENOMEM, "hipHostMalloc of %zu bytes failed", c);
                }
            }
            cap = c;
        }
        used = 0;
        return 0;
    }
    template <class T>
    Span<T> take(size_t n) {
        size_t off = (used + 255) & ~(size_t)255;
        used = off + n * sizeof(T);
        return Span<T>{(T*)(base + off), n};
    }
};

// A batch of independent jobs resident on one device (one job for the
// single-system API).  Device buffers are kept across layouts and grow only.
struct gromos_amd_system {
    int precision = 0;
    int njobs = 0;
    int32_t npad = 0, ncg = 0, ncell = 0, nmol = 0;
    std::vector<gamd::JobDev> jobs;
    std::vector<int32_t> job_n_atoms;
    int64_t bytes = 0;
    hipStream_t stream = nullptr;
    // the mixed md step's solute pass (k_gen_sol, k_excl14) runs on `side`,
    // concurrently with the solvent pair kernel on `stream` (fork / join events)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool no_side = false;   // GROMOS_AMD_SIDE_STREAM=0: everything on `stream`

    // device arrays
    gamd::JobDev* d_jobs = nullptr;
    int32_t *d_cg_start = nullptr, *d_cg_n = nullptr, *d_cg_job = nullptr, *d_atom_job = nullptr, *d_iac = nullptr;
    double *d_q = nullptr, *d_mass = nullptr, *d_inv_mass = nullptr, *d_lj = nullptr;
    int32_t *d_excl_off = nullptr, *d_excl = nullptr, *d_o14_off = nullptr, *d_o14 = nullptr;
    int32_t *d_o14s_off = nullptr, *d_o14s = nullptr;
    double *d_x[2] = {nullptr, nullptr}, *d_v[2] = {nullptr, nullptr}, *d_f = nullptr;
    double *d_ref = nullptr, *d_s_ref = nullptr;
    uint32_t *d_key_in = nullptr, *d_val_in = nullptr, *d_key_out = nullptr, *d_val_out = nullptr;
    int32_t *d_cell_count = nullptr, *d_cell_start = nullptr;
    int32_t *d_s_cg_orig = nullptr, *d_s_cg_n = nullptr, *d_s_cg_scan = nullptr, *d_s_cg_start = nullptr;
    int32_t *d_s2o = nullptr, *d_s_atom_cg = nullptr, *d_s_type = nullptr;
    int64_t *d_vv_cnt = nullptr, *d_gen_cnt = nullptr, *d_vv_off = nullptr, *d_gen_off = nullptr;
    uint32_t *d_vv_ent = nullptr, *d_gen_ent = nullptr;
    // control block of a pair-list build, read back in one copy: [0, 5) counters,
    // [5] overflow flag, [6, 6 + 2 njobs) per-job pair counts; zeroed on the device
    // at the start of each build
    uint64_t* d_ctl = nullptr;
    uint64_t* h_ctl = nullptr;   // pinned host copy
    size_t h_ctl_n = 0;
    hipEvent_t ctl_ev = nullptr;   // recorded after the control block's readback
    bool ctl_pending = false;
    uint64_t* d_counters = nullptr;
    int32_t* d_gsol = nullptr;
    int64_t *d_pc_short = nullptr, *d_pc_long = nullptr;
    int64_t* d_job_pairs = nullptr;   // 2 per job: short, long pairs of the last build
    int64_t *d_vv_cap = nullptr, *d_gen_cap = nullptr;
    int32_t *d_prev_vv = nullptr, *d_prev_gen = nullptr, *d_overflow = nullptr;
    bool have_prev = false;
    int64_t prev_tot_vv = 0, prev_tot_gen = 0;
    int64_t rebuilds = 0, refills = 0;
    double concat_ms = 0.0;   // host time concatenating job arrays in layouts
    double4* d_xq = nullptr;
    float4* d_xf = nullptr;
    long long *d_fi_s = nullptr, *d_fi_long_s = nullptr;
    int32_t nrow_blocks = 0;
    double *d_f_s = nullptr, *d_f_long_s = nullptr, *d_f_x = nullptr;
    double *d_rows = nullptr, *d_records = nullptr, *d_vv_e = nullptr, *d_vv_part = nullptr, *d_sol_e = nullptr;
    int32_t* d_sol_atom = nullptr;
    int32_t* d_excl_blocks = nullptr;   // atom blocks holding solute atoms (md-step k_excl14 grid)
    int32_t n_excl_blocks = 0;
    std::vector<int32_t> h_excl_blocks;   // host copy (the async upload's source)
    int32_t nsol = 0;
    int4* d_vv_chunk = nullptr;   // solvent pair kernel chunks (first sorted group, end, job, 0)
    int32_t* d_step = nullptr;    // md step index within the current md call
    float4* d_prec = nullptr;     // perturbed solute atoms' FP32 parameter records (2 per solute atom)
    int32_t* d_s_prec = nullptr;  // sorted atom -> record index, -1 if not perturbed
    uint32_t* d_step_done = nullptr;
    // captured plain steps (no rebuild, no outside change), one per state buffer
    // parity; replayed while the kernels' arguments are unchanged
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        gamd::SysPtrs key{};
        int64_t shape[11] = {};
    };
    StepGraph graphs[2];
    bool use_graphs = true;
    int64_t graph_launches = 0, graph_captures = 0;
    int32_t nvv_chunk = 0;
    int32_t vv_lj00 = 0;   // every job's solvent has LJ on the first-first atom pair only
    int32_t vv_neutral = 0;  // every job's solvent molecule has zero net charge
    int32_t vv_lanes = 0;  // lanes per solvent molecule in the pair kernel (0: not tuned yet)
    int64_t tuned_key = -1;
    int32_t* d_settle_fail = nullptr;
    int32_t* d_force_fail = nullptr;
    double* d_therm_scale = nullptr;
    double* d_nh_zeta = nullptr;   // per job: the Nose-Hoover chain (GROMOS_AMD_NHC_MAX each)
    double* d_nh_pend = nullptr;   // per job: the chain after the next step's update
    bool any_thermo = false;
    bool any_com = false;                  // a job removes centre-of-mass motion
    bool any_baro = false;                 // a job couples pressure
    double* d_press = nullptr;             // per job: kinetic tensor (9), then the last pressure
    int32_t* d_pg_start = nullptr;         // pressure groups of the batch (atom indices)
    int32_t* d_box_fail = nullptr;         // per job: first step whose box outgrew the cell grid
    double* d_cg_cvir = nullptr;           // per group: SETTLE's virial diagonal
    bool any_solute_shake = false;
    std::vector<int32_t> host_sh_ij;       // the batch's solute constraints (atom pairs)
    std::vector<double> host_sh_len;
    int32_t* d_sh_ij = nullptr;
    double* d_sh_len = nullptr;
    uint8_t* d_sh_skip = nullptr;
    unsigned* d_vv_sched = nullptr;        // the persistent far-pair kernel's counters
    int32_t vv_persist = 0;                // its resident blocks (0: not used)
    bool any_solvent_pass = false;         // a job constrains its solvent by SHAKE or LINCS (after the integrator)
    std::vector<double> host_lc_sdiag, host_lc_coef;   // LINCS of the solute constraints
    std::vector<int32_t> host_lc_cpl, host_lc_cidx;
    double* d_lc_sdiag = nullptr;
    int32_t* d_lc_cpl = nullptr;
    int32_t* d_lc_cidx = nullptr;
    double* d_lc_coef = nullptr;
    double* d_lc_work = nullptr;
    std::vector<int64_t> job_step0;        // per job: md steps done before the current call
    int64_t* d_job_step0 = nullptr;
    bool therm_dirty = true;   // velocities changed outside the step: rescan before the next one
    bool xq_valid = false;     // xq / xf hold the current positions in the current sorted order
    double* d_cg_ek = nullptr;
    uint8_t *d_pflag = nullptr, *d_cg_pert = nullptr;
    int32_t *d_pa_iac = nullptr, *d_pb_iac = nullptr, *d_papairs = nullptr, *d_o2s = nullptr;
    int32_t *d_pap_off = nullptr, *d_pap = nullptr;
    double *d_pa_q = nullptr, *d_pb_q = nullptr, *d_palj = nullptr, *d_pacrf = nullptr;
    int32_t n_papairs = 0;
    // covalent terms of the batch (bonded.hpp); gathered atoms = the solute atoms
    int4* d_bd_atoms = nullptr;
    double4* d_bd_prm = nullptr;
    int32_t* d_bd_kind = nullptr;
    double *d_bd_tf = nullptr, *d_bd_tev = nullptr, *d_bd_fb = nullptr;
    int32_t *d_bd_aoff = nullptr, *d_bd_aref = nullptr;
    int32_t nbd = 0;
    int32_t* d_vv_near = nullptr;   // mixed: near partners at the front of each solvent list
    std::vector<double> host_pressure;   // per job: the last md step's pressure
    std::vector<int32_t> host_pg;        // pressure-group starts of the batch (uploaded at layout)
    long long* d_fi_near = nullptr;  // mixed: the near pairs' forces (k_vv_nr)
    double* d_vv_en = nullptr;       // mixed: the near pairs' per-group records
    double near_r = gamd::VV_NEAR_R;   // their radius (nm)
    int32_t last_rec_row = -1;   // records row of the last force evaluation
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
    uint64_t host_counters[5] = {0, 0, 0, 0, 0};
    std::vector<int64_t> host_job_pairs;   // d_job_pairs of the last build
    int32_t md_step = 0;                   // step index within the current md call
    // timing of the dominant kernel
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // one timed solvent pair launch: the far (or FP64) kernel between a0 and a1;
    // mixed: the near kernel between n0 and n1; f, where set, is recorded on the
    // batch's stream before the fork that puts the near kernel on the side stream
    struct TimedLaunch {
        hipEvent_t f, a0, a1, n0, n1;
    };
    std::vector<TimedLaunch> timed;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    int64_t timed_pairs = 0;    // pair interactions the timed launches evaluated
    // per-step durations (hipEvents between steps) while enabled
    bool step_timing = false;
    std::vector<hipEvent_t> step_ev;
    std::vector<double> step_ms;

    struct Buf {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    std::unordered_map<void*, Buf> bufs;   // key: address of the member pointer

    ~gromos_amd_system();

    // Make p hold at least n elements; contents are kept only if no growth was
    // needed.  slack: grow by 25 % extra so that growing buffers settle.
    template <class T>
    int reserve(T*& p, size_t n, bool slack = false) {
        size_t need = std::max<size_t>(n, 1) * sizeof(T);
        Buf& b = bufs[(void*)&p];
        if (b.ptr && b.bytes >= need) {
            p = (T*)b.ptr;
            return 0;
        }
        size_t grown = 0;   // a buffer that outgrows its capacity: twice the need, at least double
        if (b.ptr) {
            HIPCHK(hipStreamSynchronize(stream));
            HIPCHK(hipFree(b.ptr));
            bytes -= (int64_t)b.bytes;
            grown = 2 * b.bytes;
            b.ptr = nullptr;
            b.bytes = 0;
            p = nullptr;
        }
        const size_t base = slack ? need + need / 4 + 4096 : need;
        size_t cap = slack && grown ? std::max(2 * need, grown) : base;
        if (hipMalloc(&b.ptr, cap) != hipSuccess) {
            // the doubled growth may not fit next to the rest of a full session:
            // fall back to the plain slack before giving up
            (void)hipGetLastError();
            b.ptr = nullptr;
            if (cap == base || hipMalloc(&b.ptr, base) != hipSuccess) {
                b.ptr = nullptr;
                return gamd::fail(GROMOS_AMD_ENOMEM, "hipMalloc of %zu bytes failed", base);
            }
            cap = base;
        }
        b.bytes = cap;
        bytes += (int64_t)cap;
        p = (T*)b.ptr;
        return 0;
    }
    template <class T>
    int upload(T*& p, const T* h, size_t n) {
        if (int rc = reserve(p, n, true)) return rc;
        if (n) HIPCHK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, stream));
        return 0;
    }
    template <class T>
    int upload(T*& p, const std::vector<T>& h) {
        return upload(p, h.data(), h.size());
    }
    template <class T>
    int upload(T*& p, const PinnedArena::Span<T>& h) {
        return upload(p, h.data(), h.size());
    }
    PinnedArena host;   // staging of the layout's arrays

    gamd::SysPtrs ptrs() const;
    hipEvent_t next_event();
    int collect_timing();
};

namespace gamd {

// Lay the jobs out in one batch (buffers reused, grown when needed); positions
// and velocities are zeroed, the pair list is marked stale.
int layout_batch(gromos_amd_system* S, const std::vector<const JobImage*>& imgs, int32_t precision);
int upload_jobs(gromos_amd_system* S);
// deferred: leave the single-pass fill's readback in flight (ctl_pending) for
// finish_pairlist, so that the step's first pair kernel is enqueued meanwhile
int build_pairlist(gromos_amd_system* S, bool deferred = false);
// waits for a pending readback; *refilled: the lists were refilled with exact
// slots (kernels enqueued on them since the build must run again)
int finish_pairlist(gromos_amd_system* S, bool* refilled);
int forcefield(gromos_amd_system* S, bool assemble);
// nsteps of the md sequence; records (nsteps x njobs x REC) copied to rec when
// given.  A failing job (SETTLE, or a non-finite / out-of-range force) is an
// error unless `failed` is given: it then gets each job's first failing step
// (INT32_MAX: none), and `force` (if given) whether that was a force failure.
int md_steps(gromos_amd_system* S, int32_t nsteps, std::vector<double>* rec, std::vector<int32_t>* failed = nullptr,
             std::vector<uint8_t>* force = nullptr);
int collect_step_times(gromos_amd_system* S);

// device helpers for the session
void launch_remap(double* dst, const double* src, const int32_t* map, int32_t n, hipStream_t s);
void launch_job_ke(const SysPtrs& p, double* out, int32_t njobs, hipStream_t s);
void launch_scale_v(double* v, int32_t off, int32_t n, double factor, hipStream_t s);
void launch_swap_state(double* x, double* v, int32_t a, int32_t b, int32_t n, hipStream_t s);

}  // namespace gamd
