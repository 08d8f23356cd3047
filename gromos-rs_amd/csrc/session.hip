// Sessions: continuous batching of independent MD jobs on one device.
//
// Jobs are admitted against an HBM budget (their estimated footprint is
// reserved at admission, EBUSY when it does not fit) and join the resident
// batch at the next layout boundary; finished jobs leave at one.  Layout
// boundaries fall on multiples of the shared pair-list cadence, so every job
// rebuilds its pair list on its own step 0, NSNB, 2 NSNB, ... exactly as when it
// runs alone: a job's results do not depend on the other jobs in the batch.
// A job whose step budget ends inside a segment is frozen (its state is carried
// unchanged) until it leaves.
//
// The reference runs one system per process (md.rs) and its replica exchange
// steps replicas one after another on the CPU (remd.rs:160-201); here all jobs
// step together in one set of launches.
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>

#include "host.hpp"

using namespace gamd;

namespace gamd {

__global__ void k_remap(double* dst, const double* src, const int32_t* map, int32_t n) {
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int32_t m = map[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[3 * a + c] = m >= 0 ? src[3 * (int64_t)m + c] : 0.0;
}

// 0.5 sum m v^2 of each job's current velocities, fixed order (block per job)
__global__ void __launch_bounds__(BLOCK) k_job_ke(SysPtrs p, double* out) {
    __shared__ double lds[17];
    const JobDev& J = p.jobs[blockIdx.x];
    double s = 0.0;
    for (int32_t a = J.atom_off + threadIdx.x; a < J.atom_off + J.n_atoms; a += blockDim.x) {
        const double* v = p.v_cur + 3 * (int64_t)a;
        s += 0.5 * p.mass[a] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double t[1] = {s};
    block_sum<1>(t, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = lds[16];
}

// v *= factor[job of the atom]
__global__ void k_scale_v(double* v, const int32_t* atom_job, const double* factor, int32_t npad) {
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= npad || atom_job[a] < 0) return;
    double f = factor[atom_job[a]];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[3 * a + c] *= f;
}

__global__ void k_swap_state(double* x, double* v, int32_t a, int32_t b, int32_t n3) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double t = x[3 * (int64_t)a + i];
    x[3 * (int64_t)a + i] = x[3 * (int64_t)b + i];
    x[3 * (int64_t)b + i] = t;
    t = v[3 * (int64_t)a + i];
    v[3 * (int64_t)a + i] = v[3 * (int64_t)b + i];
    v[3 * (int64_t)b + i] = t;
}

static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

void launch_remap(double* dst, const double* src, const int32_t* map, int32_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_remap, grid1(n), dim3(BLOCK), 0, s, dst, src, map, n);
}
void launch_job_ke(const SysPtrs& p, double* out, int32_t njobs, hipStream_t s) {
    if (njobs > 0) hipLaunchKernelGGL(k_job_ke, dim3(njobs), dim3(BLOCK), 0, s, p, out);
}
void launch_swap_state(double* x, double* v, int32_t a, int32_t b, int32_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_swap_state, grid1(3 * (int64_t)n), dim3(BLOCK), 0, s, x, v, a, b, 3 * n);
}

}  // namespace gamd

namespace {

// grow-only device scratch owned by the session
struct DevScratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    ~DevScratch() {
        if (ptr) (void)hipFree(ptr);
    }
    int reserve(size_t b, hipStream_t st) {
        if (b <= bytes) return 0;
        if (ptr) {
            HIPCHK(hipStreamSynchronize(st));
            HIPCHK(hipFree(ptr));
            ptr = nullptr;
            bytes = 0;
        }
        size_t cap = b + b / 4 + 4096;
        if (hipMalloc(&ptr, cap) != hipSuccess) {
            ptr = nullptr;
            return fail(GROMOS_AMD_ENOMEM, "session scratch: hipMalloc of %zu bytes failed", cap);
        }
        bytes = cap;
        return 0;
    }
};

struct SJob {
    int64_t id = 0;
    JobImage img;
    std::vector<double> pos, vel;   // host state while not resident
    std::vector<double> zeta;       // its Nose-Hoover chain across a layout (empty: at rest)
    double pressure = 0.0;          // its last md step's pressure, once it left the device
    int64_t nsteps = 0, done = 0;
    int32_t state = GROMOS_AMD_JOB_QUEUED;
    int32_t slot = -1;              // index in the resident batch, -1 if not resident
    bool evict = false;
    std::vector<double> energies;   // 5 per step done: e_lj, e_crf, e_pot, e_kin, dhdl
    std::string error;              // why the job failed
};

}  // namespace

struct gromos_amd_session {
    gromos_amd_session_params prm{};
    int64_t budget = 0;
    std::unique_ptr<gromos_amd_system> S;
    std::vector<SJob*> batch;
    std::map<int64_t, std::unique_ptr<SJob>> jobs;
    std::deque<SJob*> queue;
    int64_t next_id = 1, reserved = 0, clock = 0, layouts = 0, pair_count = 0;
    int64_t reserved_entries = 0;   // estimated neighbour-list entries of queued and resident jobs
    double layout_ms = 0.0;
    double concat_ms_gone = 0.0;   // concatenation time of batches already released
    DevScratch stage, map, small;
    bool step_timing = false;
    // per-launch timing of the solvent pair kernel, kept across batch changes
    bool kernel_timing = false;
    double kern_ms_gone = 0.0;
    int64_t kern_launches_gone = 0, kern_pairs_gone = 0;
    std::vector<double> step_ms;
};

static int set_device(gromos_amd_session* s) {
    HIPCHK(hipSetDevice(s->prm.device));
    return 0;
}

static int running_count(const gromos_amd_session* s) {
    int n = 0;
    for (const SJob* j : s->batch)
        if (j->state == GROMOS_AMD_JOB_RUNNING && !j->evict) ++n;
    return n;
}

// a batch system about to be released: its counters move to the session's
static void retire_timing(gromos_amd_session* s, gromos_amd_system* S) {
    s->concat_ms_gone += S->concat_ms;
    if (S->collect_timing() != 0) (void)hipGetLastError();
    s->kern_ms_gone += S->timed_ms;
    s->kern_launches_gone += S->timed_launches;
    s->kern_pairs_gone += S->timed_pairs;
}

// a job's admission reservations end
static void release(gromos_amd_session* s, SJob* j) {
    s->reserved -= j->img.footprint;
    s->reserved_entries -= j->img.list_entries;
}

// Change the resident batch: jobs done or evicted leave (done jobs' final
// state is copied to the host), queued jobs join.  Continuing jobs' positions
// and velocities move to their new offsets on the device.  If the new layout
// fails (device memory), the continuing jobs stop as FAILED with their last
// state copied to the host, the queued jobs stay queued, and the device batch
// is released; the error is returned.
static int relayout(gromos_amd_session* s) {
    auto t0 = std::chrono::steady_clock::now();
    gromos_amd_system* S = s->S.get();
    std::vector<SJob*> keep, leave;
    for (SJob* j : s->batch)
        (j->state == GROMOS_AMD_JOB_DONE || j->state == GROMOS_AMD_JOB_FAILED || j->evict ? leave : keep).push_back(j);
    std::vector<SJob*> nb = keep;
    for (SJob* j : s->queue) nb.push_back(j);
    std::vector<int32_t> old_off;
    int32_t old_npad = S ? S->npad : 0;
    if (S) {
        hipStream_t st = S->stream;
        for (SJob* j : leave) {
            if (j->evict) continue;
            if (j->img.J.baro) {
                rebox_image(j->img, S->jobs[j->slot].box);
                j->pressure = S->host_pressure[j->slot];
            }
            size_t n3 = (size_t)j->img.n * 3, off = (size_t)S->jobs[j->slot].atom_off * 3;
            j->pos.resize(n3);
            j->vel.resize(n3);
            HIPCHK(hipMemcpyAsync(j->pos.data(), S->d_x[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(j->vel.data(), S->d_v[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        if (!keep.empty()) {
            // the continuing jobs' Nose-Hoover chains (layout_batch starts every
            // chain at zero; they are put back below)
            std::vector<double> z((size_t)GROMOS_AMD_NHC_MAX * S->njobs);
            HIPCHK(hipMemcpyAsync(z.data(), S->d_nh_zeta, z.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (SJob* j : keep) {
                j->zeta.assign(z.begin() + (size_t)GROMOS_AMD_NHC_MAX * j->slot,
                               z.begin() + (size_t)GROMOS_AMD_NHC_MAX * (j->slot + 1));
                // a barostat's box as it is now (the new layout sizes its cells from it)
                if (j->img.J.baro) rebox_image(j->img, S->jobs[j->slot].box);
            }
            size_t n3 = (size_t)old_npad * 3;
            if (int rc = s->stage.reserve(2 * n3 * sizeof(double), st)) return rc;
            double* sx = (double*)s->stage.ptr;
            HIPCHK(hipMemcpyAsync(sx, S->d_x[S->cur], n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(sx + n3, S->d_v[S->cur], n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
            for (SJob* j : keep) old_off.push_back(S->jobs[j->slot].atom_off);
        }
        HIPCHK(hipStreamSynchronize(st));
    }
    for (SJob* j : leave) {
        j->slot = -1;
        release(s, j);
        if (j->evict) s->jobs.erase(j->id);
    }
    std::deque<SJob*> queued = s->queue;
    s->queue.clear();
    s->batch.clear();
    if (nb.empty()) {
        if (S) retire_timing(s, S);
        s->S.reset();
        s->layouts++;
        s->layout_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    }
    if (!S) {
        s->S = std::make_unique<gromos_amd_system>();
        S = s->S.get();
        S->step_timing = s->step_timing;
        S->timing = s->kernel_timing;
    }
    std::vector<const JobImage*> imgs;
    int64_t batch_atoms = 0;
    for (SJob* j : nb) {
        imgs.push_back(&j->img);
        batch_atoms += j->img.n;
    }
    // GROMOS_AMD_LAYOUT_ATOM_LIMIT (a test hook): a layout of more atoms fails
    // as if device memory ran out, before anything is allocated
    const char* lim = std::getenv("GROMOS_AMD_LAYOUT_ATOM_LIMIT");
    const bool injected = lim && *lim && batch_atoms > std::atoll(lim);
    if (int rc = injected ? fail(GROMOS_AMD_ENOMEM, "layout of %lld atoms over GROMOS_AMD_LAYOUT_ATOM_LIMIT",
                                 (long long)batch_atoms)
                          : layout_batch(S, imgs, s->prm.precision)) {
        const std::string why = gromos_amd_last_error();
        // the continuing jobs' state is in the session's staging copy
        const double* sx = keep.empty() ? nullptr : (const double*)s->stage.ptr;
        for (size_t k = 0; k < keep.size(); ++k) {
            SJob* j = keep[k];
            size_t n3 = (size_t)j->img.n * 3, off = (size_t)old_off[k] * 3;
            j->pos.resize(n3);
            j->vel.resize(n3);
            if (hipMemcpy(j->pos.data(), sx + off, n3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(j->vel.data(), sx + (size_t)old_npad * 3 + off, n3 * sizeof(double),
                          hipMemcpyDeviceToHost) != hipSuccess)
                (void)hipGetLastError();
            j->slot = -1;
            j->state = GROMOS_AMD_JOB_FAILED;
            j->error = "batch layout failed: " + why;
            release(s, j);
        }
        s->queue = queued;
        if (S) retire_timing(s, S);
        s->S.reset();
        s->layouts++;
        return fail(rc, "%s", why.c_str());
    }
    hipStream_t st = S->stream;
    if (!keep.empty()) {
        std::vector<int32_t> mp(S->npad, -1);
        for (size_t k = 0; k < keep.size(); ++k) {
            int32_t no = S->jobs[k].atom_off, oo = old_off[k];
            for (int32_t a = 0; a < keep[k]->img.n; ++a) mp[no + a] = oo + a;
        }
        if (int rc = s->map.reserve(mp.size() * sizeof(int32_t), st)) return rc;
        HIPCHK(hipMemcpyAsync(s->map.ptr, mp.data(), mp.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        const double* sx = (const double*)s->stage.ptr;
        launch_remap(S->d_x[0], sx, (const int32_t*)s->map.ptr, S->npad, st);
        launch_remap(S->d_v[0], sx + (size_t)old_npad * 3, (const int32_t*)s->map.ptr, S->npad, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));   // mp feeds an async copy
    }
    for (size_t k = 0; k < keep.size(); ++k)
        if (!keep[k]->zeta.empty())
            HIPCHK(hipMemcpyAsync(S->d_nh_zeta + (size_t)GROMOS_AMD_NHC_MAX * k, keep[k]->zeta.data(),
                                  GROMOS_AMD_NHC_MAX * sizeof(double), hipMemcpyHostToDevice, st));
    for (size_t k = keep.size(); k < nb.size(); ++k) {
        SJob* j = nb[k];
        size_t n3 = (size_t)j->img.n * 3, off = (size_t)S->jobs[k].atom_off * 3;
        HIPCHK(hipMemcpyAsync(S->d_x[0] + off, j->pos.data(), n3 * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(S->d_v[0] + off, j->vel.data(), n3 * sizeof(double), hipMemcpyHostToDevice, st));
        j->state = GROMOS_AMD_JOB_RUNNING;
    }
    HIPCHK(hipStreamSynchronize(st));
    for (size_t k = keep.size(); k < nb.size(); ++k) {
        std::vector<double>().swap(nb[k]->pos);
        std::vector<double>().swap(nb[k]->vel);
    }
    for (size_t k = 0; k < nb.size(); ++k) nb[k]->slot = (int32_t)k;
    s->batch = nb;
    s->layouts++;
    s->layout_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

static int boundary(gromos_amd_session* s) {
    bool any_leave = false;
    for (SJob* j : s->batch)
        any_leave = any_leave || j->state == GROMOS_AMD_JOB_DONE || j->state == GROMOS_AMD_JOB_FAILED || j->evict;
    if (!any_leave && s->queue.empty()) return 0;
    bool idle = running_count(s) == 0;
    if (!idle && s->clock % s->prm.admit_interval != 0) return 0;
    return relayout(s);
}

extern "C" int gromos_amd_session_create(const gromos_amd_session_params* p, gromos_amd_session** out) {
    if (!p || !out) return fail(GROMOS_AMD_EINVAL, "session_create: null argument");
    *out = nullptr;
    if (p->precision != GROMOS_AMD_FP64 && p->precision != GROMOS_AMD_MIXED)
        return fail(GROMOS_AMD_EINVAL, "precision must be 0 (FP64) or 1 (mixed)");
    if (p->nsnb < 1) return fail(GROMOS_AMD_EINVAL, "nsnb must be >= 1");
    if (p->admit_interval < 0 || (p->admit_interval > 0 && p->admit_interval % p->nsnb != 0))
        return fail(GROMOS_AMD_EINVAL, "admit_interval must be a multiple of nsnb");
    auto s = std::make_unique<gromos_amd_session>();
    s->prm = *p;
    if (s->prm.admit_interval == 0) s->prm.admit_interval = p->nsnb;
    if (int rc = set_device(s.get())) return rc;
    if (p->hbm_budget > 0) s->budget = p->hbm_budget;
    else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        s->budget = (int64_t)(0.9 * (double)fr);
    }
    *out = s.release();
    return 0;
}

extern "C" void gromos_amd_session_destroy(gromos_amd_session* s) {
    if (!s) return;
    (void)hipSetDevice(s->prm.device);
    delete s;
}

extern "C" int64_t gromos_amd_job_footprint(const gromos_amd_topology* t, const gromos_amd_params* p) {
    if (!t || !p) return fail(GROMOS_AMD_EINVAL, "job_footprint: null argument");
    JobImage im;
    if (int rc = make_image(*t, *p, 0, im)) return rc;
    return im.footprint;
}

extern "C" int gromos_amd_job_admit(gromos_amd_session* s, const gromos_amd_topology* t, const gromos_amd_params* p,
                                    const double* pos, const double* vel, int64_t nsteps, int64_t* job_id) {
    if (!s || !t || !p || !pos || !job_id || nsteps < 0) return fail(GROMOS_AMD_EINVAL, "job_admit: bad arguments");
    if (p->nsnb != s->prm.nsnb) return fail(GROMOS_AMD_EINVAL, "job_admit: nsnb %d differs from the session's %d", p->nsnb, s->prm.nsnb);
    auto j = std::make_unique<SJob>();
    if (int rc = make_image(*t, *p, 0, j->img)) return rc;
    for (const SJob* o : s->batch)
        if (o->img.solvent && j->img.solvent && o->img.J.aps != j->img.J.aps)
            return fail(GROMOS_AMD_EINVAL, "job_admit: solvent molecule size differs from the resident jobs'");
    for (const SJob* o : s->queue)
        if (o->img.solvent && j->img.solvent && o->img.J.aps != j->img.J.aps)
            return fail(GROMOS_AMD_EINVAL, "job_admit: solvent molecule size differs from the queued jobs'");
    if (j->img.footprint > s->budget)
        return fail(GROMOS_AMD_ENOMEM, "job_admit: footprint %lld B exceeds the session budget %lld B",
                    (long long)j->img.footprint, (long long)s->budget);
    if (s->reserved + j->img.footprint > s->budget)
        return fail(GROMOS_AMD_EBUSY, "job_admit: %lld B reserved of %lld B, job needs %lld B", (long long)s->reserved,
                    (long long)s->budget, (long long)j->img.footprint);
    // the batch's neighbour lists are indexed with 32-bit offsets
    if (j->img.list_entries >= MAX_LIST_ENTRIES)
        return fail(GROMOS_AMD_ENOMEM, "job_admit: about %lld neighbour-list entries exceed the 32-bit list offsets",
                    (long long)j->img.list_entries);
    if (s->reserved_entries + j->img.list_entries >= MAX_LIST_ENTRIES)
        return fail(GROMOS_AMD_EBUSY, "job_admit: about %lld neighbour-list entries reserved, job needs %lld (limit %lld)",
                    (long long)s->reserved_entries, (long long)j->img.list_entries, (long long)MAX_LIST_ENTRIES);
    size_t n3 = (size_t)t->n_atoms * 3;
    j->pos.assign(pos, pos + n3);
    if (vel) j->vel.assign(vel, vel + n3);
    else j->vel.assign(n3, 0.0);
    j->nsteps = nsteps;
    j->id = s->next_id++;
    if (nsteps == 0) j->state = GROMOS_AMD_JOB_DONE;
    else s->queue.push_back(j.get());
    if (nsteps > 0) {
        s->reserved += j->img.footprint;
        s->reserved_entries += j->img.list_entries;
    }
    *job_id = j->id;
    s->jobs[j->id] = std::move(j);
    return 0;
}

extern "C" int gromos_amd_session_step(gromos_amd_session* s, int32_t nsteps) {
    if (!s || nsteps < 0) return fail(GROMOS_AMD_EINVAL, "session_step: bad arguments");
    if (int rc = set_device(s)) return rc;
    const int32_t nsnb = s->prm.nsnb;
    std::vector<double> rec;
    while (nsteps > 0) {
        if (s->clock % nsnb == 0)
            if (int rc = boundary(s)) return rc;
        // one md call up to the next admission boundary (a job that finishes on the
        // way is held by its stop step); while jobs wait in the queue, no further
        // than the first NSNB boundary at which every resident job has finished
        // (an idle batch takes them in at once)
        const int32_t ai = s->prm.admit_interval;
        int32_t span = ai - (int32_t)(s->clock % ai);
        if (!s->queue.empty()) {
            int64_t idle_in = 0;   // steps until no resident job runs
            for (const SJob* j : s->batch)
                if (j->state == GROMOS_AMD_JOB_RUNNING && !j->evict) idle_in = std::max<int64_t>(idle_in, j->nsteps - j->done);
            const int64_t to_idle = ((s->clock + idle_in + nsnb - 1) / nsnb) * nsnb - s->clock;
            span = (int32_t)std::min<int64_t>(span, std::max<int64_t>(to_idle, nsnb - s->clock % nsnb));
        }
        int32_t k = std::min<int32_t>(nsteps, span);
        gromos_amd_system* S = s->S.get();
        if (S && running_count(s) > 0) {
            bool changed = false;
            std::vector<int64_t> adv(s->batch.size(), 0);
            for (size_t b = 0; b < s->batch.size(); ++b) {
                SJob* j = s->batch[b];
                int64_t rem = (j->state == GROMOS_AMD_JOB_RUNNING && !j->evict) ? j->nsteps - j->done : 0;
                adv[b] = std::min<int64_t>(rem, k);
                int32_t stop = rem >= k ? INT32_MAX : (int32_t)rem;
                if (S->jobs[b].stop != stop) {
                    S->jobs[b].stop = stop;
                    changed = true;
                }
            }
            if (changed)
                if (int rc = upload_jobs(S)) return rc;
            for (size_t b = 0; b < s->batch.size(); ++b) S->job_step0[b] = s->batch[b]->done;
            int64_t pc0 = S->pair_count;
            std::vector<int32_t> failed;
            std::vector<uint8_t> fkind;
            if (int rc = md_steps(S, k, &rec, &failed, &fkind)) return rc;
            s->pair_count += S->pair_count - pc0;
            if (s->step_timing) {
                s->step_ms.insert(s->step_ms.end(), S->step_ms.begin(), S->step_ms.end());
                S->step_ms.clear();
            }
            for (size_t b = 0; b < s->batch.size(); ++b) {
                SJob* j = s->batch[b];
                if (failed[b] != INT32_MAX && failed[b] < adv[b]) {
                    // this job alone stops; the others are independent and go on
                    adv[b] = failed[b];
                    j->state = GROMOS_AMD_JOB_FAILED;
                    char msg[160];
                    if (fkind[b] == 2)
                        snprintf(msg, sizeof(msg), "the barostat shrank the box below the pair search's cells (step %lld)",
                                 (long long)(j->done + failed[b]));
                    else if (fkind[b])
                        snprintf(msg, sizeof(msg), "non-finite or out-of-range force (|F| >= %.3g kJ/mol/nm) (step %lld)",
                                 FORCE_LIMIT, (long long)(j->done + failed[b]));
                    else
                        snprintf(msg, sizeof(msg), "constraints failed: SETTLE sin(angle) > 1.0 or SHAKE unconverged after 1000 sweeps (step %lld)",
                                 (long long)(j->done + failed[b]));
                    j->error = msg;
                }
                for (int64_t st = 0; st < adv[b]; ++st) {
                    const double* r = rec.data() + ((size_t)st * S->njobs + b) * REC;
                    j->energies.insert(j->energies.end(), r, r + 5);
                }
                j->done += adv[b];
                if (j->state == GROMOS_AMD_JOB_RUNNING && j->done >= j->nsteps) j->state = GROMOS_AMD_JOB_DONE;
            }
        }
        s->clock += k;
        nsteps -= k;
    }
    // a job that finished on the last boundary leaves now if the batch is idle
    if (s->clock % nsnb == 0 && running_count(s) == 0)
        if (int rc = boundary(s)) return rc;
    return 0;
}

static SJob* find(gromos_amd_session* s, int64_t id) {
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        fail(GROMOS_AMD_ENOTFOUND, "no job %lld in this session", (long long)id);
        return nullptr;
    }
    return it->second.get();
}

extern "C" int gromos_amd_job_query(gromos_amd_session* s, int64_t id, int32_t* state, int64_t* steps_done) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "job_query: null session");
    SJob* j = find(s, id);
    if (!j) return GROMOS_AMD_ENOTFOUND;
    if (state) *state = j->state;
    if (steps_done) *steps_done = j->done;
    if (j->state == GROMOS_AMD_JOB_FAILED) fail(GROMOS_AMD_ESETTLE, "%s", j->error.c_str());
    return 0;
}

extern "C" int gromos_amd_job_box(gromos_amd_session* s, int64_t id, double* box, double* pressure) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "job_box: null session");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(GROMOS_AMD_EINVAL, "job_box: no job %lld", (long long)id);
    const SJob* j = it->second.get();
    const gromos_amd_system* S = s->S.get();
    const bool resident = S && j->slot >= 0;
    const Box& b = resident ? S->jobs[j->slot].box : j->img.J.box;
    if (box)
        for (int a = 0; a < 3; ++a) box[a] = b.L[a];
    if (pressure) *pressure = resident ? S->host_pressure[j->slot] : j->pressure;
    return 0;
}

extern "C" int gromos_amd_job_fetch(gromos_amd_session* s, int64_t id, double* pos, double* vel) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "job_fetch: null session");
    if (int rc = set_device(s)) return rc;
    SJob* j = find(s, id);
    if (!j) return GROMOS_AMD_ENOTFOUND;
    size_t n3 = (size_t)j->img.n * 3;
    if (j->slot >= 0) {
        gromos_amd_system* S = s->S.get();
        size_t off = (size_t)S->jobs[j->slot].atom_off * 3;
        if (pos) HIPCHK(hipMemcpyAsync(pos, S->d_x[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        if (vel) HIPCHK(hipMemcpyAsync(vel, S->d_v[S->cur] + off, n3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        HIPCHK(hipStreamSynchronize(S->stream));
    } else {
        if (pos) std::memcpy(pos, j->pos.data(), n3 * sizeof(double));
        if (vel) std::memcpy(vel, j->vel.data(), n3 * sizeof(double));
    }
    return 0;
}

extern "C" int gromos_amd_job_energies(gromos_amd_session* s, int64_t id, int64_t first, int64_t count, double* out) {
    if (!s || !out || first < 0 || count < 0) return fail(GROMOS_AMD_EINVAL, "job_energies: bad arguments");
    SJob* j = find(s, id);
    if (!j) return GROMOS_AMD_ENOTFOUND;
    if (first + count > j->done) return fail(GROMOS_AMD_EINVAL, "job_energies: steps [%lld, %lld) not done (%lld done)",
                                             (long long)first, (long long)(first + count), (long long)j->done);
    std::memcpy(out, j->energies.data() + 5 * first, (size_t)count * 5 * sizeof(double));
    return 0;
}

extern "C" int gromos_amd_job_evict(gromos_amd_session* s, int64_t id) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "job_evict: null session");
    SJob* j = find(s, id);
    if (!j) return GROMOS_AMD_ENOTFOUND;
    if (j->slot >= 0) {   // leaves the device at the next layout boundary
        j->evict = true;
        return 0;
    }
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
        if (*it == j) {
            s->queue.erase(it);
            release(s, j);
            break;
        }
    s->jobs.erase(id);
    return 0;
}

// resident jobs only: kinetic energy, velocity scaling and configuration
// exchange act on the device state between steps
static int resident_slots(gromos_amd_session* s, int32_t n, const int64_t* ids, std::vector<int32_t>& slots) {
    slots.resize(n);
    for (int32_t k = 0; k < n; ++k) {
        SJob* j = find(s, ids[k]);
        if (!j) return GROMOS_AMD_ENOTFOUND;
        if (j->slot < 0) return fail(GROMOS_AMD_EINVAL, "job %lld is not resident", (long long)ids[k]);
        slots[k] = j->slot;
    }
    return 0;
}

extern "C" int gromos_amd_jobs_kinetic_energy(gromos_amd_session* s, int32_t n, const int64_t* ids, double* out) {
    if (!s || n < 0 || (n && (!ids || !out))) return fail(GROMOS_AMD_EINVAL, "jobs_kinetic_energy: bad arguments");
    if (int rc = set_device(s)) return rc;
    std::vector<int32_t> slots;
    if (int rc = resident_slots(s, n, ids, slots)) return rc;
    if (n == 0) return 0;
    gromos_amd_system* S = s->S.get();
    if (int rc = s->small.reserve((size_t)S->njobs * sizeof(double), S->stream)) return rc;
    launch_job_ke(S->ptrs(), (double*)s->small.ptr, S->njobs, S->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> ke(S->njobs);
    HIPCHK(hipMemcpyAsync(ke.data(), s->small.ptr, ke.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    for (int32_t k = 0; k < n; ++k) out[k] = ke[slots[k]];
    return 0;
}

extern "C" int gromos_amd_jobs_scale_velocities(gromos_amd_session* s, int32_t n, const int64_t* ids,
                                                const double* factors) {
    if (!s || n < 0 || (n && (!ids || !factors))) return fail(GROMOS_AMD_EINVAL, "jobs_scale_velocities: bad arguments");
    if (int rc = set_device(s)) return rc;
    std::vector<int32_t> slots;
    if (int rc = resident_slots(s, n, ids, slots)) return rc;
    if (n == 0) return 0;
    gromos_amd_system* S = s->S.get();
    std::vector<double> f(S->njobs, 1.0);
    for (int32_t k = 0; k < n; ++k) f[slots[k]] = factors[k];
    if (int rc = s->small.reserve(f.size() * sizeof(double), S->stream)) return rc;
    HIPCHK(hipMemcpyAsync(s->small.ptr, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, S->stream));
    hipLaunchKernelGGL(k_scale_v, dim3((S->npad + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, S->stream, S->d_v[S->cur],
                       S->d_atom_job, (const double*)s->small.ptr, S->npad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S->stream));
    S->therm_dirty = true;
    return 0;
}

extern "C" int gromos_amd_jobs_swap_configurations(gromos_amd_session* s, int64_t a, int64_t b) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "jobs_swap_configurations: null session");
    if (int rc = set_device(s)) return rc;
    int64_t ids[2] = {a, b};
    std::vector<int32_t> slots;
    if (int rc = resident_slots(s, 2, ids, slots)) return rc;
    if (s->clock % s->prm.nsnb != 0)
        return fail(GROMOS_AMD_EINVAL, "jobs_swap_configurations: only between steps on a pair-list boundary (clock %lld)",
                    (long long)s->clock);
    gromos_amd_system* S = s->S.get();
    SJob* ja = s->batch[slots[0]];
    SJob* jb = s->batch[slots[1]];
    if (ja->img.n != jb->img.n) return fail(GROMOS_AMD_EINVAL, "jobs_swap_configurations: atom counts differ");
    if (slots[0] == slots[1]) return 0;
    launch_swap_state(S->d_x[S->cur], S->d_v[S->cur], S->jobs[slots[0]].atom_off, S->jobs[slots[1]].atom_off,
                      ja->img.n, S->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S->stream));
    S->built = false;   // the swapped jobs' lists belong to the other configuration
    S->long_computed = false;
    S->therm_dirty = true;
    S->xq_valid = false;
    return 0;
}

extern "C" int gromos_amd_session_get_info(gromos_amd_session* s, gromos_amd_session_info* o) {
    if (!s || !o) return fail(GROMOS_AMD_EINVAL, "session_get_info: null argument");
    std::memset(o, 0, sizeof(*o));
    for (auto& kv : s->jobs) {
        const SJob* j = kv.second.get();
        if (j->state == GROMOS_AMD_JOB_QUEUED) o->queued++;
        else if (j->state == GROMOS_AMD_JOB_RUNNING) o->running++;
        else o->done++;
    }
    o->resident_jobs = (int32_t)s->batch.size();
    o->device_bytes = s->S ? s->S->bytes : 0;
    o->reserved_bytes = s->reserved;
    o->budget = s->budget;
    o->clock = s->clock;
    o->layouts = s->layouts;
    o->layout_ms = s->layout_ms;
    o->layout_concat_ms = s->concat_ms_gone + (s->S ? s->S->concat_ms : 0.0);
    o->pair_count = s->pair_count;
    return 0;
}

extern "C" int gromos_amd_session_step_times(gromos_amd_session* s, int32_t enable, double* ms, int64_t cap,
                                             int64_t* n) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "session_step_times: null session");
    if (n) *n = (int64_t)s->step_ms.size();
    if (ms)
        for (int64_t k = 0; k < std::min<int64_t>(cap, (int64_t)s->step_ms.size()); ++k) ms[k] = s->step_ms[k];
    if (enable >= 0) {
        s->step_timing = enable != 0;
        if (s->step_timing) s->step_ms.clear();
        if (s->S) s->S->step_timing = s->step_timing;
    }
    return 0;
}

extern "C" int gromos_amd_session_kernel_timing(gromos_amd_session* s, int32_t enable, double* total_ms,
                                                int64_t* launches, int64_t* pairs) {
    if (!s) return fail(GROMOS_AMD_EINVAL, "session_kernel_timing: null session");
    if (int rc = set_device(s)) return rc;
    gromos_amd_system* S = s->S.get();
    if (S)
        if (int rc = S->collect_timing()) return rc;
    if (total_ms) *total_ms = s->kern_ms_gone + (S ? S->timed_ms : 0.0);
    if (launches) *launches = s->kern_launches_gone + (S ? S->timed_launches : 0);
    if (pairs) *pairs = s->kern_pairs_gone + (S ? S->timed_pairs : 0);
    if (enable >= 0) {
        const bool on = enable != 0;
        if (on && !s->kernel_timing) {
            s->kern_ms_gone = 0.0;
            s->kern_launches_gone = 0;
            s->kern_pairs_gone = 0;
            if (S) {
                S->timed_ms = 0.0;
                S->timed_launches = 0;
                S->timed_pairs = 0;
            }
        }
        s->kernel_timing = on;
        if (S) S->timing = on;
    }
    return 0;
}
