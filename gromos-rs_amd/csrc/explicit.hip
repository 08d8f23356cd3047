// Pair-list kernels: GPU drop-ins for the gromos-forces free functions that
// take an explicit pair list (crates/gromos-forces/src/nonbonded/
// innerloops.rs:318-733, excluded.rs:25-206, perturbed.rs:309-386).
//
// One lane per listed pair (or per group / sentinel / atom).  Forces go to the
// caller's f64 buffer with hardware FP64 atomics.  Energies and the virial are
// summed per block in a fixed order into a scratch row per block, and the rows
// are folded in a fixed order by one more block (k_fold), so they do not
// depend on scheduling.
#include <mutex>

#include "device_common.hpp"
#include "explicit.hpp"

namespace gamd {

__device__ __forceinline__ void atomic_add3(double* f, int64_t i, double x, double y, double z) {
    atomicAdd(f + 3 * i, x);
    atomicAdd(f + 3 * i + 1, y);
    atomicAdd(f + 3 * i + 2, z);
}

// the block's sums of v (e_lj, e_crf, virial[9]) into its scratch row
__device__ __forceinline__ void flush_block(double (&v)[EXPL_ROW], double* part) {
    __shared__ double lds[17 * EXPL_ROW];
    block_sum<EXPL_ROW>(v, lds);
    if (threadIdx.x < EXPL_ROW) part[EXPL_ROW * (int64_t)blockIdx.x + threadIdx.x] = lds[16 * EXPL_ROW + threadIdx.x];
}

// one block: the nrows scratch rows summed in a fixed order (lane t takes rows
// t, t + 256, ..., then the block sum) and added to energies / virial
__global__ void __launch_bounds__(EXPL_BLOCK) k_fold(const double* part, int64_t nrows, double* energies,
                                                     double* virial) {
    __shared__ double lds[17 * EXPL_ROW];
    double v[EXPL_ROW] = {0};
    for (int64_t r = threadIdx.x; r < nrows; r += EXPL_BLOCK)
#pragma unroll
        for (int c = 0; c < EXPL_ROW; ++c) v[c] += part[EXPL_ROW * r + c];
    block_sum<EXPL_ROW>(v, lds);
    if (threadIdx.x < 2) atomicAdd(energies + threadIdx.x, lds[16 * EXPL_ROW + threadIdx.x]);
    if (virial && threadIdx.x >= 2 && threadIdx.x < EXPL_ROW)
        atomicAdd(virial + threadIdx.x - 2, lds[16 * EXPL_ROW + threadIdx.x]);
}

__device__ __forceinline__ void acc_pair(double (&v)[11], d3 r, double fm, double elj, double ecrf,
                                         bool vir, double* forces, int64_t i, int64_t j) {
    double fx = r.x * fm, fy = r.y * fm, fz = r.z * fm;
    atomic_add3(forces, i, fx, fy, fz);
    atomic_add3(forces, j, -fx, -fy, -fz);
    v[0] += elj;
    v[1] += ecrf;
    if (vir) {
        double rr[3] = {r.x, r.y, r.z}, ff[3] = {fx, fy, fz};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) v[2 + 3 * a + b] += rr[a] * ff[b];
    }
}

// process_pairs (innerloops.rs:169-210)
__global__ void __launch_bounds__(EXPL_BLOCK) k_plain(ExplArgs a) {
    double v[11] = {0};
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.n) {
        int64_t i = a.pairs[2 * k], j = a.pairs[2 * k + 1];
        d3 r = nim(a.box, ld3(a.pos, i), ld3(a.pos, j));
        const double* p = a.lj + 4 * ((int64_t)a.iac[i] * a.n_types + a.iac[j]);
        double qp = a.q[i] * a.q[j] * a.fpepsi;
        double fm, elj, ecrf;
        lj_crf64(r.x, r.y, r.z, p[0], p[1], qp, a.crf, fm, elj, ecrf);
        acc_pair(v, r, fm, elj, ecrf, a.virial != nullptr, a.forces, i, j);
    }
    flush_block(v, a.part);
}

// process_pairs_cg_grouped (innerloops.rs:105-162): one lane per group
__global__ void __launch_bounds__(EXPL_BLOCK) k_cg_grouped(ExplArgs a) {
    double v[11] = {0};
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.n) {
        const uint32_t* G = a.groups + 4 * g;
        d3 ri = ld3(a.pos, G[0]), rj = ld3(a.pos, G[1]);
        d3 rr = nim(a.box, ri, rj);
        double tx = rr.x - ri.x + rj.x, ty = rr.y - ri.y + rj.y, tz = rr.z - ri.z + rj.z;
        for (uint32_t k = G[2]; k < G[3]; ++k) {
            int64_t i = a.pairs[2 * k], j = a.pairs[2 * k + 1];
            d3 pi = ld3(a.pos, i), pj = ld3(a.pos, j);
            d3 r{pi.x + tx - pj.x, pi.y + ty - pj.y, pi.z + tz - pj.z};
            const double* p = a.lj + 4 * ((int64_t)a.iac[i] * a.n_types + a.iac[j]);
            double qp = a.q[i] * a.q[j] * a.fpepsi;
            double fm, elj, ecrf;
            lj_crf64(r.x, r.y, r.z, p[0], p[1], qp, a.crf, fm, elj, ecrf);
            acc_pair(v, r, fm, elj, ecrf, a.virial != nullptr, a.forces, i, j);
        }
    }
    flush_block(v, a.part);
}

// process_solvent_pairs (innerloops.rs:218-309): tables from the first pair's molecule
__global__ void __launch_bounds__(EXPL_BLOCK) k_solvent(ExplArgs a) {
    double v[11] = {0};
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.aps;
    if (k < a.n) {
        const int64_t first = a.pairs[0];
        int64_t i0 = a.pairs[2 * k], j0 = a.pairs[2 * k + 1];
        d3 pi0 = ld3(a.pos, i0), pj0 = ld3(a.pos, j0);
        d3 rf = nim(a.box, pi0, pj0);
        double tx = rf.x - pi0.x + pj0.x, ty = rf.y - pi0.y + pj0.y, tz = rf.z - pi0.z + pj0.z;
        for (int ai = 0; ai < n; ++ai) {
            d3 pi = ld3(a.pos, i0 + ai);
            double xi = pi.x + tx, yi = pi.y + ty, zi = pi.z + tz;
            double qi = a.q[first + ai];
            for (int aj = 0; aj < n; ++aj) {
                d3 pj = ld3(a.pos, j0 + aj);
                d3 r{xi - pj.x, yi - pj.y, zi - pj.z};
                const double* p = a.lj + 4 * ((int64_t)a.iac[first + ai] * a.n_types + a.iac[first + aj]);
                double qp = qi * a.q[first + aj] * a.fpepsi;
                double fm, elj, ecrf;
                lj_crf64(r.x, r.y, r.z, p[0], p[1], qp, a.crf, fm, elj, ecrf);
                acc_pair(v, r, fm, elj, ecrf, a.virial != nullptr, a.forces, i0 + ai, j0 + aj);
            }
        }
    }
    flush_block(v, a.part);
}

// rf_excluded_interactions (excluded.rs:25-112): one lane per atom i, pairs j > i
__global__ void __launch_bounds__(EXPL_BLOCK) k_rf_excluded(ExclArgs a) {
    double v[11] = {0};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_atoms) {
        double qi = a.q[i];
        d3 pi = ld3(a.pos, i);
        if (i < a.n_solute) {
            v[1] += -0.5 * qi * qi * a.fpepsi * a.crf.crf_cut;
            for (int k = a.off[i]; k < a.off[i + 1]; ++k) {
                int64_t j = a.idx[k];
                if (j <= i) continue;
                d3 r = nim(a.box, pi, ld3(a.pos, j));
                double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
                double qp = qi * a.q[j] * a.fpepsi;
                v[1] += qp * (-a.crf.crf_2cut3i * r2 - a.crf.crf_cut);
                double fs = qp * a.crf.crf_cut3i;
                double fx = r.x * fs, fy = r.y * fs, fz = r.z * fs;
                atomic_add3(a.forces, i, fx, fy, fz);
                atomic_add3(a.forces, j, -fx, -fy, -fz);
                double rr[3] = {r.x, r.y, r.z}, ff[3] = {fx, fy, fz};
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) v[2 + 3 * x + y] += rr[x] * ff[y];
            }
        } else {
            for (int k = a.off[i]; k < a.off[i + 1]; ++k) {
                int64_t j = a.idx[k];
                if (j <= i) continue;
                d3 r = nim(a.box, pi, ld3(a.pos, j));
                double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
                v[1] += -qi * a.q[j] * a.fpepsi * a.crf.crf_2cut3i * r2;
            }
        }
    }
    flush_block(v, a.part);
}

// one_four_interaction_loop (excluded.rs:131-206)
__global__ void __launch_bounds__(EXPL_BLOCK) k_one_four(ExclArgs a) {
    double v[11] = {0};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_atoms) {
        d3 pi = ld3(a.pos, i);
        for (int k = a.off[i]; k < a.off[i + 1]; ++k) {
            int64_t j = a.idx[k];
            if (j <= i) continue;
            d3 r = nim(a.box, pi, ld3(a.pos, j));
            double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            if (r2 < 1e-10) continue;
            const double* p = a.lj + 4 * ((int64_t)a.iac[i] * a.n_types + a.iac[j]);
            double inv_r2 = 1.0 / r2;
            double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            double elj = (p[3] * inv_r6 - p[2]) * inv_r6;
            double flj = (12.0 * p[3] * inv_r6 - 6.0 * p[2]) * inv_r6 * inv_r2;
            double qp = a.q[i] * a.q[j] * a.fpepsi;
            double inv_r = sqrt(inv_r2);
            double ecrf = qp * (inv_r * a.coulomb_scaling - a.crf.crf_2cut3i * r2 - a.crf.crf_cut);
            double fcrf = qp * (inv_r * a.coulomb_scaling * inv_r2 + a.crf.crf_cut3i);
            double ft = flj + fcrf;
            double fx = r.x * ft, fy = r.y * ft, fz = r.z * ft;
            atomic_add3(a.forces, i, fx, fy, fz);
            atomic_add3(a.forces, j, -fx, -fy, -fz);
            v[0] += elj;
            v[1] += ecrf;
            double rr[3] = {r.x, r.y, r.z}, ff[3] = {fx, fy, fz};
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) v[2 + 3 * x + y] += rr[x] * ff[y];
        }
    }
    flush_block(v, a.part);
}

// perturbed_pairlist_correction (perturbed.rs:309-386)
__global__ void __launch_bounds__(EXPL_BLOCK) k_pert_pairs(PertArgs a) {
    __shared__ double lds[17 * 3];
    double v[3] = {0, 0, 0};
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.n) {
        int64_t i = a.pairs[2 * k], j = a.pairs[2 * k + 1];
        if (a.flag[i]) {
            bool pj = a.flag[j] != 0;
            int a_iac_i = a.a_iac[i], b_iac_i = a.b_iac[i];
            int a_iac_j = pj ? a.a_iac[j] : (int)a.iac[j], b_iac_j = pj ? a.b_iac[j] : (int)a.iac[j];
            double a_q = a.a_q[i] * (pj ? a.a_q[j] : a.q[j]);
            double b_q = a.b_q[i] * (pj ? a.b_q[j] : a.q[j]);
            double alj, acrf;
            if (pj) {
                alj = (a.alpha_lj[i] + a.alpha_lj[j]) * 0.5;
                acrf = (a.alpha_crf[i] + a.alpha_crf[j]) * 0.5;
            } else {
                alj = a.alpha_lj[i];
                acrf = a.alpha_crf[i];
            }
            const double* la = a.lj + 4 * ((int64_t)a_iac_i * a.n_types + a_iac_j);
            const double* lb = a.lj + 4 * ((int64_t)b_iac_i * a.n_types + b_iac_j);
            d3 r = nim(a.box, ld3(a.pos, i), ld3(a.pos, j));
            double fa, elja, ecrfa, o[5];
            lj_crf64(r.x, r.y, r.z, la[0], la[1], a_q * a.fpepsi, a.crf, fa, elja, ecrfa);
            pert_lj_crf64(r.x, r.y, r.z, la[0], la[1], lb[0], lb[1], a_q, b_q, alj, acrf, a.lp,
                          a.crf, a.fpepsi, o);
            double d = o[0] - fa;
            double fx = r.x * d, fy = r.y * d, fz = r.z * d;
            atomic_add3(a.forces, i, fx, fy, fz);
            atomic_add3(a.forces, j, -fx, -fy, -fz);
            v[0] = o[1] - elja;
            v[1] = o[2] - ecrfa;
            v[2] = o[3] + o[4];
        }
    }
    block_sum<3>(v, lds);
    if (threadIdx.x < 3) atomicAdd(a.out + threadIdx.x, lds[16 * 3 + threadIdx.x]);
}

// dual-topology parameters of the pair (i, j), either atom perturbed
// (perturbed.rs:458-470, 531-562): a perturbed atom takes its A / B type and
// charge, an unperturbed one its regular ones in both states; alpha is the
// perturbed atom's, or the mean of both when both are perturbed.
struct PairAB {
    const double* la;
    const double* lb;
    double a_q, b_q, alj, acrf;
};
__device__ __forceinline__ PairAB pair_ab(const PertArgs& a, int64_t i, int64_t j, bool pi, bool pj) {
    PairAB q{nullptr, nullptr, 0.0, 0.0, 0.0, 0.0};
    if (a.lj) {  // the excluded pass has no LJ part (and no type arrays)
        int a_iac_i = pi ? a.a_iac[i] : (int)a.iac[i], b_iac_i = pi ? a.b_iac[i] : (int)a.iac[i];
        int a_iac_j = pj ? a.a_iac[j] : (int)a.iac[j], b_iac_j = pj ? a.b_iac[j] : (int)a.iac[j];
        q.la = a.lj + 4 * ((int64_t)a_iac_i * a.n_types + a_iac_j);
        q.lb = a.lj + 4 * ((int64_t)b_iac_i * a.n_types + b_iac_j);
    }
    q.a_q = (pi ? a.a_q[i] : a.q[i]) * (pj ? a.a_q[j] : a.q[j]);
    q.b_q = (pi ? a.b_q[i] : a.q[i]) * (pj ? a.b_q[j] : a.q[j]);
    if (pi && pj) {
        q.alj = (a.alpha_lj[i] + a.alpha_lj[j]) * 0.5;
        q.acrf = (a.alpha_crf[i] + a.alpha_crf[j]) * 0.5;
    } else {
        int64_t w = pi ? i : j;
        q.alj = a.alpha_lj[w];
        q.acrf = a.alpha_crf[w];
    }
    return q;
}

__device__ __forceinline__ void flush3(double (&v)[3], double* out, int nv) {
    __shared__ double lds[17 * 3];
    block_sum<3>(v, lds);
    if ((int)threadIdx.x < nv) atomicAdd(out + threadIdx.x, lds[16 * 3 + threadIdx.x]);
}

// perturbed_self_energy_correction (perturbed.rs:395-425): one lane per atom
__global__ void __launch_bounds__(EXPL_BLOCK) k_pert_self(PertArgs a) {
    double v[3] = {0, 0, 0};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_atoms && a.flag[i]) pert_self64(a.a_q[i], a.b_q[i], a.lp, a.crf, a.fpepsi, v[0], v[1]);
    flush3(v, a.out, 2);
}

// perturbed_excluded_correction (perturbed.rs:431-511): one lane per solute
// atom i, its excluded partners j > i with at least one of the two perturbed
__global__ void __launch_bounds__(EXPL_BLOCK) k_pert_excluded(PertArgs a) {
    double v[3] = {0, 0, 0};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_solute) {
        const bool pi = a.flag[i] != 0;
        d3 xi = ld3(a.pos, i);
        for (int32_t k = a.off[i]; k < a.off[i + 1]; ++k) {
            int64_t j = a.idx[k];
            if (j <= i) continue;
            const bool pj = a.flag[j] != 0;
            if (!pi && !pj) continue;
            PairAB q = pair_ab(a, i, j, pi, pj);
            d3 r = nim(a.box, xi, ld3(a.pos, j));
            double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            PertDelta d = pert_excl64(r2, q.a_q, q.b_q, q.acrf, a.lp, a.crf, a.fpepsi);
            double fx = r.x * d.f, fy = r.y * d.f, fz = r.z * d.f;
            atomic_add3(a.forces, i, fx, fy, fz);
            atomic_add3(a.forces, j, -fx, -fy, -fz);
            v[1] += d.de_crf;
            v[2] += d.dhdl;
        }
    }
    flush3(v, a.out, 3);
}

// perturbed_one_four_correction (perturbed.rs:517-605): one lane per atom i,
// its 1-4 partners j > i with at least one of the two perturbed
__global__ void __launch_bounds__(EXPL_BLOCK) k_pert_one_four(PertArgs a) {
    double v[3] = {0, 0, 0};
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_atoms) {
        const bool pi = a.flag[i] != 0;
        d3 xi = ld3(a.pos, i);
        for (int32_t k = a.off[i]; k < a.off[i + 1]; ++k) {
            int64_t j = a.idx[k];
            if (j <= i) continue;
            const bool pj = a.flag[j] != 0;
            if (!pi && !pj) continue;
            PairAB q = pair_ab(a, i, j, pi, pj);
            d3 r = nim(a.box, xi, ld3(a.pos, j));
            double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            if (r2 < 1e-10) continue;
            PertDelta d = pert_14_64(r.x, r.y, r.z, r2, q.la, q.lb, q.a_q, q.b_q, q.alj, q.acrf, a.lp, a.crf,
                                     a.fpepsi);
            double fx = r.x * d.f, fy = r.y * d.f, fz = r.z * d.f;
            atomic_add3(a.forces, i, fx, fy, fz);
            atomic_add3(a.forces, j, -fx, -fy, -fz);
            v[0] += d.de_lj;
            v[1] += d.de_crf;
            v[2] += d.dhdl;
        }
    }
    flush3(v, a.out, 3);
}

// perturbed_atom_pair_correction (perturbed.rs:624-707): one lane per
// PERTATOMPAIR entry; entries that are not 1-4 pairs are skipped
__global__ void __launch_bounds__(EXPL_BLOCK) k_pert_atom_pairs(PertArgs a) {
    double v[3] = {0, 0, 0};
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < a.n) {
        const int32_t* ap = a.atom_pairs + 4 * k;
        int64_t i = ap[0], j = ap[1];
        bool in14 = false;
        for (int32_t t = a.off[i]; t < a.off[i + 1]; ++t) in14 |= a.idx[t] == j;
        for (int32_t t = a.off[j]; t < a.off[j + 1]; ++t) in14 |= a.idx[t] == i;
        if (in14) {
            d3 r = nim(a.box, ld3(a.pos, i), ld3(a.pos, j));
            double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            if (r2 >= 1e-10) {
                const double* lj = a.lj + 4 * ((int64_t)a.iac[i] * a.n_types + a.iac[j]);
                PertDelta d = pert_atom_pair64(r2, lj, a.q[i] * a.q[j] * a.fpepsi, ap[2], ap[3], a.lp, a.crf);
                double fx = r.x * d.f, fy = r.y * d.f, fz = r.z * d.f;
                atomic_add3(a.forces, i, fx, fy, fz);
                atomic_add3(a.forces, j, -fx, -fy, -fz);
                v[0] = d.de_lj;
                v[1] = d.de_crf;
                v[2] = d.dhdl;
            }
        }
    }
    flush3(v, a.out, 3);
}

// perturbed_lj_crf_interaction (perturbed.rs:124-266) on independent pairs
__global__ void __launch_bounds__(EXPL_BLOCK) k_soft_pairs(SoftPairArgs a) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const double* r = a.r + 3 * k;
    const double* q = a.prm + 8 * k;
    double o[5];
    pert_lj_crf64(r[0], r[1], r[2], q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], a.lp, a.crf, a.fpepsi, o);
#pragma unroll
    for (int c = 0; c < 5; ++c) a.out[5 * k + c] = o[c];
}

static int blocks_for(int64_t n) { return (int)expl_blocks(n); }

// kernel K over n lanes with its partial rows in part, then the fold
template <class Args>
static hipError_t launch_folded(void (*K)(Args), const Args& a, int64_t n, double* part, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(K, dim3(blocks_for(n)), dim3(EXPL_BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(EXPL_BLOCK), 0, s, (const double*)part, expl_blocks(n), a.energies,
                       a.virial);
    return hipGetLastError();
}

hipError_t launch_plain(const ExplArgs& a, hipStream_t s) { return launch_folded(k_plain, a, a.n, a.part, s); }
hipError_t launch_cg_grouped(const ExplArgs& a, hipStream_t s) {
    return launch_folded(k_cg_grouped, a, a.n, a.part, s);
}
hipError_t launch_solvent(const ExplArgs& a, hipStream_t s) { return launch_folded(k_solvent, a, a.n, a.part, s); }
hipError_t launch_rf_excluded(const ExclArgs& a, hipStream_t s) {
    return launch_folded(k_rf_excluded, a, (int64_t)a.n_atoms, a.part, s);
}
hipError_t launch_one_four(const ExclArgs& a, hipStream_t s) {
    return launch_folded(k_one_four, a, (int64_t)a.n_atoms, a.part, s);
}

// The no-allocation variant's partial rows: one static array per device,
// reached from host code only through the address hipGetSymbolAddress returns
// for the current device (the array's name in host code is not a device
// address).  Every call on a device shares it: calls that overlap there race.
static __device__ __align__(256) double scratch[NOALLOC_BLOCKS * EXPL_ROW];
constexpr int MAX_DEVICES = 64;
static double* g_scratch[MAX_DEVICES];
static std::mutex g_scratch_mu;

static hipError_t scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

hipError_t launch_plain_noalloc(ExplArgs a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.n > NOALLOC_MAX_PAIRS) return hipErrorInvalidValue;   // the caller reports the limit
    hipError_t e = scratch_ptr(&a.part);
    if (e != hipSuccess) return e;
    return launch_folded(k_plain, a, a.n, a.part, s);
}
hipError_t launch_pert_pairs(const PertArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pert_pairs, dim3(blocks_for(a.n)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pert_self(const PertArgs& a, hipStream_t s) {
    if (a.n_atoms <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pert_self, dim3(blocks_for(a.n_atoms)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_pert_excluded(const PertArgs& a, hipStream_t s) {
    if (a.n_solute <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pert_excluded, dim3(blocks_for(a.n_solute)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_pert_one_four(const PertArgs& a, hipStream_t s) {
    if (a.n_atoms <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pert_one_four, dim3(blocks_for(a.n_atoms)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_pert_atom_pairs(const PertArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pert_atom_pairs, dim3(blocks_for(a.n)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_soft_pairs(const SoftPairArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soft_pairs, dim3(blocks_for(a.n)), dim3(EXPL_BLOCK), 0, s, a);
    return hipGetLastError();
}

}  // namespace gamd
