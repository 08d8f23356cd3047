// Covalent-term kernels and the standalone calculate_bonded_forces_ntf entry
// (crates/gromos-forces/src/bonded/mod.rs:141-185).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/gromos_amd.h"
#include "bonded.hpp"
#include "host.hpp"

namespace gamd {

namespace {

constexpr int BD_BLOCK = 256;

__global__ void __launch_bounds__(BD_BLOCK) k_bd_terms(BondedDev bd, const double* __restrict__ x) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bd.nterms) return;
    const int4 a = bd.atoms[t];
    const double4 p = bd.prm[t];
    bd::TermOut o;
    bd::zero(o);
    switch (bd.kind[t]) {
        case BD_BOND: bd::bond(ld3(x, a.x), ld3(x, a.y), p, o); break;
        case BD_ANGLE: bd::angle(ld3(x, a.x), ld3(x, a.y), ld3(x, a.z), p, o); break;
        case BD_DIHEDRAL: bd::dihedral(ld3(x, a.x), ld3(x, a.y), ld3(x, a.z), ld3(x, a.w), p, o); break;
        default: bd::improper(ld3(x, a.x), ld3(x, a.y), ld3(x, a.z), ld3(x, a.w), p, o); break;
    }
    double* tf = bd.tf + 12 * (int64_t)t;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        tf[3 * s] = o.f[s].x;
        tf[3 * s + 1] = o.f[s].y;
        tf[3 * s + 2] = o.f[s].z;
    }
    double* ev = bd.tev + BD_EV * (int64_t)t;
    ev[0] = o.e;
#pragma unroll
    for (int c = 0; c < 9; ++c) ev[1 + c] = o.vir[c];
}

// per atom: its contributions in term order; a kind's sum starts from zero and
// is added to the atom's total when the kind changes (the reference's per-kind
// buffers, then ForceEnergy::add in the order bonds, angles, dihedrals, impropers)
__global__ void __launch_bounds__(BD_BLOCK) k_bd_gather(BondedDev bd) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= bd.natoms) return;
    double f[3] = {0.0, 0.0, 0.0}, part[3] = {0.0, 0.0, 0.0};
    const int32_t e0 = bd.aoff[w], e1 = bd.aoff[w + 1];
    int32_t kprev = -1;
    for (int32_t e = e0; e < e1; ++e) {
        const int32_t r = bd.aref[e];
        const int32_t t = r >> 2, s = r & 3;
        const int32_t k = bd.kind[t];
        if (k != kprev) {
            if (kprev >= 0)
                for (int c = 0; c < 3; ++c) {
                    f[c] += part[c];
                    part[c] = 0.0;
                }
            kprev = k;
        }
        const double* tf = bd.tf + 12 * (int64_t)t + 3 * s;
        for (int c = 0; c < 3; ++c) part[c] += tf[c];
    }
    if (kprev >= 0)
        for (int c = 0; c < 3; ++c) f[c] += part[c];
    for (int c = 0; c < 3; ++c) bd.fb[3 * (int64_t)w + c] = f[c];
}

__global__ void __launch_bounds__(BD_BLOCK) k_bd_add(BondedDev bd, double* f) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= bd.natoms) return;
    const int64_t o = bd.amap ? bd.amap[w] : w;
    for (int c = 0; c < 3; ++c) f[3 * o + c] = bd.fb[3 * (int64_t)w + c] + f[3 * o + c];
}

// standalone entry: per kind energy and the virial summed over all terms
// (one block, fixed order): out = e_bond, e_angle, e_dihedral, e_improper, virial[9]
__global__ void __launch_bounds__(BD_BLOCK) k_bd_sum(BondedDev bd, double* out) {
    constexpr int W = 13;
    __shared__ double lds[17 * W];
    double v[W] = {0};
    for (int32_t t = threadIdx.x; t < bd.nterms; t += blockDim.x) {
        const double* ev = bd.tev + BD_EV * (int64_t)t;
        const int k = bd.kind[t];
        for (int q = 0; q < 4; ++q) v[q] += (q == k) ? ev[0] : 0.0;
        for (int c = 0; c < 9; ++c) v[4 + c] += ev[1 + c];
    }
    block_sum<W>(v, lds);
    if (threadIdx.x < W) out[threadIdx.x] = lds[16 * W + threadIdx.x];
}

}  // namespace

void launch_bonded(const BondedDev& bd, const double* x, hipStream_t s) {
    if (bd.nterms > 0) k_bd_terms<<<(bd.nterms + BD_BLOCK - 1) / BD_BLOCK, BD_BLOCK, 0, s>>>(bd, x);
    if (bd.natoms > 0) k_bd_gather<<<(bd.natoms + BD_BLOCK - 1) / BD_BLOCK, BD_BLOCK, 0, s>>>(bd);
}

void launch_bonded_add(const BondedDev& bd, double* f, hipStream_t s) {
    if (bd.natoms > 0) k_bd_add<<<(bd.natoms + BD_BLOCK - 1) / BD_BLOCK, BD_BLOCK, 0, s>>>(bd, f);
}

// ---------------------------------------------------------------------------
// host: terms of one topology
// ---------------------------------------------------------------------------

int bonded_terms(const gromos_amd_bonded& b, int32_t n_atoms, int job, BondedTerms& out) {
    out = BondedTerms{};
    struct Kind {
        int32_t n, width, ntypes;
        const int32_t* idx;
        const double* types;
        int32_t type_width;
        bool on;
        const char* name;
    };
    // NTF (md.rs:331-335): ntf[0] bonds, ntf[1] angles, ntf[2] impropers, ntf[3] dihedrals
    const Kind kinds[4] = {
        {b.n_bonds, 2, b.n_bond_types, b.bonds, b.bond_types, 3, b.ntf[0] != 0, "bond"},
        {b.n_angles, 3, b.n_angle_types, b.angles, b.angle_types, 3, b.ntf[1] != 0, "angle"},
        {b.n_dihedrals, 4, b.n_dihedral_types, b.dihedrals, b.dihedral_types, 3, b.ntf[3] != 0, "dihedral"},
        {b.n_impropers, 4, b.n_improper_types, b.impropers, b.improper_types, 2, b.ntf[2] != 0, "improper"},
    };
    for (int k = 0; k < 4; ++k) {
        const Kind& K = kinds[k];
        if (K.n < 0 || K.ntypes < 0) return fail(GROMOS_AMD_EINVAL, "job %d: negative %s count", job, K.name);
        if (K.n > 0 && !K.idx) return fail(GROMOS_AMD_EINVAL, "job %d: %s indices missing", job, K.name);
        if (K.ntypes > 0 && !K.types) return fail(GROMOS_AMD_EINVAL, "job %d: %s types missing", job, K.name);
        out.n[k] = 0;
        if (!K.on) continue;   // the kind is switched off: no term of it enters
        const int32_t rw = K.width + 1;
        for (int32_t t = 0; t < K.n; ++t) {
            const int32_t* e = K.idx + (int64_t)rw * t;
            int4 a = make_int4(-1, -1, -1, -1);
            int32_t* ap = &a.x;
            for (int s = 0; s < K.width; ++s) {
                if (e[s] < 0 || e[s] >= n_atoms)
                    return fail(GROMOS_AMD_EINVAL, "job %d: %s %d atom index out of range", job, K.name, t);
                ap[s] = e[s];
            }
            const int32_t ty = e[K.width];
            if (ty < 0) return fail(GROMOS_AMD_EINVAL, "job %d: %s %d has a negative type", job, K.name, t);
            if (ty >= K.ntypes) continue;   // bonds.rs:16-18 and the like: skipped
            const double* p = K.types + (int64_t)K.type_width * ty;
            double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
            switch (k) {
                case BD_BOND: q.x = p[0]; q.y = p[2]; break;                 // k_quartic, r0
                case BD_ANGLE: q.x = p[0]; q.y = std::cos(p[2]); break;      // k_cosine, theta0.cos()
                case BD_DIHEDRAL: {
                    const double m = p[2];
                    if (m != std::floor(m))
                        return fail(GROMOS_AMD_EINVAL, "job %d: dihedral type %d multiplicity not an integer", job, ty);
                    if (m < 0.0 || m > 6.0) continue;   // unsupported multiplicity: skipped (dihedrals.rs:84-87)
                    q.x = p[0]; q.y = p[1]; q.z = m;
                    break;
                }
                default: q.x = p[0]; q.y = p[1]; break;                     // K, q0
            }
            out.atoms.push_back(a);
            out.prm.push_back(q);
            out.kind.push_back(k);
            out.n[k]++;
        }
    }
    return 0;
}

// CSR of the gathered atoms' contributions (entries term << 2 | slot in term
// order); atom ids local to [0, natoms)
void bonded_csr(const BondedTerms& t, int32_t natoms, std::vector<int32_t>& aoff, std::vector<int32_t>& aref) {
    aoff.assign((size_t)natoms + 1, 0);
    const int32_t nt = (int32_t)t.atoms.size();
    for (int32_t k = 0; k < nt; ++k) {
        const int32_t* ap = &t.atoms[k].x;
        for (int s = 0; s < 4; ++s)
            if (ap[s] >= 0) aoff[ap[s] + 1]++;
    }
    for (int32_t a = 0; a < natoms; ++a) aoff[a + 1] += aoff[a];
    aref.assign((size_t)aoff[natoms], 0);
    std::vector<int32_t> at(aoff.begin(), aoff.end() - 1);
    for (int32_t k = 0; k < nt; ++k) {
        const int32_t* ap = &t.atoms[k].x;
        for (int s = 0; s < 4; ++s)
            if (ap[s] >= 0) aref[at[ap[s]]++] = (k << 2) | s;
    }
}

}  // namespace gamd

using namespace gamd;

// ---------------------------------------------------------------------------
// standalone plan
// ---------------------------------------------------------------------------

struct gromos_amd_bonded_plan {
    int32_t n_atoms = 0;
    int32_t nterms = 0;
    int4* atoms = nullptr;
    double4* prm = nullptr;
    int32_t* kind = nullptr;
    double* tf = nullptr;
    double* tev = nullptr;
    int32_t* aoff = nullptr;
    int32_t* aref = nullptr;
    int32_t n[4] = {0, 0, 0, 0};
    ~gromos_amd_bonded_plan() {
        for (void* p : {(void*)atoms, (void*)prm, (void*)kind, (void*)tf, (void*)tev, (void*)aoff, (void*)aref})
            if (p) (void)hipFree(p);
    }
};

template <class T>
static int dev_copy(T*& d, const std::vector<T>& h) {
    const size_t b = std::max<size_t>(h.size(), 1) * sizeof(T);
    HIPCHK(hipMalloc(&d, b));
    if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int gromos_amd_bonded_plan_create(const gromos_amd_bonded* b, gromos_amd_bonded_plan** out) {
    if (!b || !out) return fail(GROMOS_AMD_EINVAL, "bonded_plan_create: bad arguments");
    *out = nullptr;
    if (b->n_atoms <= 0) return fail(GROMOS_AMD_EINVAL, "bonded_plan_create: n_atoms must be > 0");
    BondedTerms t;
    if (int rc = bonded_terms(*b, b->n_atoms, 0, t)) return rc;
    if (t.atoms.size() >= (size_t)(INT32_MAX >> 2)) return fail(GROMOS_AMD_EINVAL, "bonded_plan_create: too many terms");
    std::vector<int32_t> aoff, aref;
    bonded_csr(t, b->n_atoms, aoff, aref);
    auto* P = new gromos_amd_bonded_plan();
    P->n_atoms = b->n_atoms;
    P->nterms = (int32_t)t.atoms.size();
    std::memcpy(P->n, t.n, sizeof P->n);
    int rc = dev_copy(P->atoms, t.atoms);
    rc = rc ? rc : dev_copy(P->prm, t.prm);
    rc = rc ? rc : dev_copy(P->kind, t.kind);
    rc = rc ? rc : dev_copy(P->aoff, aoff);
    rc = rc ? rc : dev_copy(P->aref, aref);
    if (!rc && hipMalloc(&P->tf, std::max<size_t>(P->nterms, 1) * 12 * sizeof(double)) != hipSuccess)
        rc = fail(GROMOS_AMD_ENOMEM, "bonded_plan_create: device memory");
    if (!rc && hipMalloc(&P->tev, std::max<size_t>(P->nterms, 1) * BD_EV * sizeof(double)) != hipSuccess)
        rc = fail(GROMOS_AMD_ENOMEM, "bonded_plan_create: device memory");
    if (rc) {
        delete P;
        return rc;
    }
    *out = P;
    return 0;
}

extern "C" void gromos_amd_bonded_plan_destroy(gromos_amd_bonded_plan* p) { delete p; }

extern "C" int gromos_amd_bonded_plan_terms(const gromos_amd_bonded_plan* p, int32_t counts[4]) {
    if (!p || !counts) return fail(GROMOS_AMD_EINVAL, "bonded_plan_terms: bad arguments");
    std::memcpy(counts, p->n, 4 * sizeof(int32_t));
    return 0;
}

extern "C" int gromos_amd_bonded_forces(const gromos_amd_bonded_plan* p, const double* pos, double* forces,
                                        double* out, void* stream) {
    if (!p || !pos || !forces || !out) return fail(GROMOS_AMD_EINVAL, "bonded_forces: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    BondedDev bd{};
    bd.atoms = p->atoms;
    bd.prm = p->prm;
    bd.kind = p->kind;
    bd.tf = p->tf;
    bd.tev = p->tev;
    bd.aoff = p->aoff;
    bd.aref = p->aref;
    bd.amap = nullptr;
    bd.fb = forces;
    bd.nterms = p->nterms;
    bd.natoms = p->n_atoms;
    launch_bonded(bd, pos, s);
    k_bd_sum<<<1, BD_BLOCK, 0, s>>>(bd, out);
    HIPCHK(hipGetLastError());
    return 0;
}
