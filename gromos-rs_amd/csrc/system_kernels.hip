// Kernels of the device-resident path (plain FP64 and mixed-precision forms).
//
//  pair list      k_cg_ref / k_scatter / k_cell_order / k_sorted_atoms / k_caps /
//                 k_nbl<FILL> / k_totals / k_job_pairs
//                                    StandardPairlistAlgorithm semantics
//                                    (pairlist.rs:183-442) on a cell grid
//  forces         k_gather, k_vv / k_vv_mx, k_gen / k_gen_sv / k_gen_mx / k_gen_sol,
//                 k_excl14, k_assemble
//                                    (innerloops.rs:105-309, excluded.rs:25-206,
//                                     forcefield.rs:420-756)
//  step           k_therm, k_integrate (leap-frog, SETTLE, E_kin), k_vv_parts,
//                 k_reduce           (leap_frog.rs:37-148, constraints.rs:815-1036,
//                                     temperature.rs:63-87)
#include <algorithm>

#include "system_kernels.hpp"

namespace gamd {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// With the blocks of one XCD renumbered to a contiguous range of the (cell-
// sorted, hence spatially ordered) work, each L2 holds one slab of positions
// and its halo instead of the whole system.  The grid is a multiple of NXCD;
// returns -1 for padding blocks (whole blocks exit, no block barrier is hit).
constexpr int NXCD = 8;
static_assert(NXCD_SCHED == NXCD + 1, "one counter per XCD and a finished-wave count");
__device__ __forceinline__ int32_t xcd_block(int32_t nreal) {
    const int32_t per = (int32_t)gridDim.x / NXCD;
    int32_t b = ((int32_t)blockIdx.x % NXCD) * per + (int32_t)blockIdx.x / NXCD;
    return b < nreal ? b : -1;
}
static inline dim3 xcd_grid(int64_t nblocks) {
    return dim3((unsigned)((nblocks + NXCD - 1) / NXCD * NXCD));
}

// ---------------------------------------------------------------------------
// pair list
// ---------------------------------------------------------------------------

// Reference point of each charge group: COG for solute groups (left-to-right sum
// from zero, then division, topology.rs:408-416), first atom for solvent
// (pairlist.rs:304-307, 359-367).  Cell key from the wrapped point.
// Also clears the build's control block (nctl words: counters, overflow flag,
// per-job pair counts) and counts the groups per cell (cell_count is zero here:
// cleared at layout and counted back to zero by the previous build's k_scatter).
__global__ void k_cg_ref(SysPtrs p, int32_t ncg, int32_t nctl) {
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncg) return;
    for (int32_t w = c; w < nctl; w += ncg) p.counters[w] = 0ull;
    const JobDev& J = p.jobs[p.cg_job[c]];
    int32_t s = p.cg_start[c], e = s + p.cg_n[c];
    const double* x = p.x_cur;
    double rx, ry, rz;
    if (c - J.cg_off < J.n_solute_cg) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int32_t a = s; a < e; ++a) {
            sx += x[3 * a];
            sy += x[3 * a + 1];
            sz += x[3 * a + 2];
        }
        double n = (double)(e - s);
        rx = sx / n;
        ry = sy / n;
        rz = sz / n;
    } else {
        rx = x[3 * s];
        ry = x[3 * s + 1];
        rz = x[3 * s + 2];
    }
    p.ref[3 * c] = rx;
    p.ref[3 * c + 1] = ry;
    p.ref[3 * c + 2] = rz;
    int32_t cc[3] = {0, 0, 0};   // vacuum and triclinic jobs: one cell
    if (J.box.kind == 1) {
        double r[3] = {rx, ry, rz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double u = r[a] * J.box.invL[a];
            u = u - floor(u);
            int32_t k = (int32_t)(u * (double)J.ncell[a]);
            cc[a] = k < 0 ? 0 : (k >= J.ncell[a] ? J.ncell[a] - 1 : k);
        }
    }
    const uint32_t key = (uint32_t)(J.cell_off + cc[0] + J.ncell[0] * (cc[1] + J.ncell[1] * cc[2]));
    p.key_in[c] = key;
    atomicAdd(p.cell_count + key, 1);
}

// Counting sort of the groups by cell key, stable (ascending original index
// within a cell, so a job's order does not depend on the batch).  k_scatter
// drops each group into its cell's range at an arbitrary position (counting
// cell_count back to zero for the next build); k_cell_order then ranks each
// cell's members by original index, a wave per cell, and writes the sorted
// group arrays: original group, atom count, reference point.
__global__ void k_scatter(SysPtrs p, int32_t ncg) {
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncg) return;
    const uint32_t key = p.key_in[c];
    const int32_t r = atomicSub(p.cell_count + key, 1) - 1;
    p.val_out[p.cell_start[key] + r] = (uint32_t)c;
}

__global__ void __launch_bounds__(BLOCK) k_cell_order(SysPtrs p, int32_t ncell) {
    const int32_t cell = (int32_t)blockIdx.x * (BLOCK / WAVE) + (int32_t)(threadIdx.x / WAVE);
    if (cell >= ncell) return;   // wave-uniform
    const int lane = threadIdx.x & (WAVE - 1);
    const int32_t s = p.cell_start[cell], n = p.cell_start[cell + 1] - s;
    auto put = [&](int32_t k, int32_t c) {
        p.s_cg_orig[k] = c;
        p.s_cg_n[k] = p.cg_n[c];
        p.s_ref[3 * k] = p.ref[3 * c];
        p.s_ref[3 * k + 1] = p.ref[3 * c + 1];
        p.s_ref[3 * k + 2] = p.ref[3 * c + 2];
    };
    if (n <= WAVE) {
        const int32_t v = lane < n ? (int32_t)p.val_out[s + lane] : INT32_MAX;
        int32_t rank = 0;
        for (int32_t j = 0; j < n; ++j) rank += __shfl(v, j, WAVE) < v ? 1 : 0;
        if (lane < n) put(s + rank, v);
    } else {
        // large cells (few cells per axis, or no periodic box): quadratic ranks
        for (int32_t i = lane; i < n; i += WAVE) {
            const int32_t v = (int32_t)p.val_out[s + i];
            int32_t rank = 0;
            for (int32_t j = 0; j < n; ++j) rank += (int32_t)p.val_out[s + j] < v ? 1 : 0;
            put(s + rank, v);
        }
    }
}

// Sorted atom layout: job j's groups fill [atom_off, atom_off + n_atoms) in
// sorted order.  s_cg_start comes from an exclusive scan of s_cg_n, shifted per
// job to the job's aligned offset.  Also writes the sorted position copies xq
// (and, mixed, xf: FP32 offset from the group's first atom, difference taken in
// FP64) of the current positions, as k_gather does.
__global__ void k_sorted_atoms(SysPtrs p, int32_t ncg, int32_t mixed) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncg) return;
    int32_t c = p.s_cg_orig[k];
    const JobDev& J = p.jobs[p.cg_job[c]];
    // groups of one job are contiguous in sorted space (the key carries cell_off)
    int32_t start = p.s_cg_scan[k] - p.s_cg_scan[J.cg_off] + J.atom_off;
    p.s_cg_start[k] = start;
    int32_t n = p.s_cg_n[k];
    int32_t o = p.cg_start[c];
    for (int32_t a = 0; a < n; ++a) {
        p.s2o[start + a] = o + a;
        p.o2s[o + a] = start + a;
        p.s_atom_cg[start + a] = k;
        p.s_type[start + a] = p.iac[o + a];
        p.s_prec[start + a] = (J.has_pert && p.pflag[o + a]) ? J.sol_off + (o + a - J.atom_off) : -1;
        const double4 v = make_double4(p.x_cur[3 * (o + a)], p.x_cur[3 * (o + a) + 1], p.x_cur[3 * (o + a) + 2],
                                       p.q[o + a]);
        p.xq[start + a] = v;
        if (mixed)
            p.xf[start + a] = make_float4((float)(v.x - p.x_cur[3 * o]), (float)(v.y - p.x_cur[3 * o + 1]),
                                          (float)(v.z - p.x_cur[3 * o + 2]), (float)v.w);
    }
}

// canonical squared distance of two reference points: nearest_image(lower, higher)
__device__ __forceinline__ double ref_d2(const Box& b, d3 lo, d3 hi) {
    d3 r = nim_exact(b, lo, hi);
    return r.x * r.x + r.y * r.y + r.z * r.z;
}

// Neighbour search, one block per cell.  The groups of a cell share its
// stencil (the cells within two along each axis, all cells along an axis with
// fewer than five), so the block stages the stencil's candidate groups in LDS,
// NBL_PIECE at a time (coalesced loads by all four waves), and each wave tests
// them against its share of the cell's groups (group s0 + w + 4 t in lane t of
// wave w), the group's data broadcast from the lane that holds it.  Survivors
// are compacted per group with a ballot, so each group's entries are written
// contiguously and in a fixed order (cells in z-y-x order, candidates in
// sorted order).  Classification as the reference (pairlist.rs:252-415).
// nearest_image is exactly antisymmetric (IEEE subtraction and half-away
// rounding are), so the squared distance is the reference's whichever group
// comes first.
// FILL = false: count only.  FILL = true: write at most cap[k] entries per
// group, record the true counts and raise *overflow when a group had more.
// Pair counts: the cell's total goes to the cell's first group, zeros to the
// others (only their sum is used).
constexpr int CELL_R = 2;        // neighbour cells searched along each axis
constexpr int NBL_RUNS = 2 * (2 * CELL_R + 1) * (2 * CELL_R + 1);   // x runs of the stencil
constexpr int NBL_PIECE = 512;   // candidates staged in LDS at a time
constexpr int NBL_WAVES = BLOCK / WAVE;

__device__ __forceinline__ int32_t bcast(int32_t v, int32_t t) { return __builtin_amdgcn_readlane(v, t); }
__device__ __forceinline__ double bcast(double v, int32_t t) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)b, t);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(b >> 32), t);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// nearest_image for the search, the same vector as nim(): rint (one instruction)
// equals the reference's half-away rounding except at exact halves (where
// x - rint(x), exact, is +-0.5), which take round() in a branch that the wave
// skips unless one of its lanes has one
__device__ __forceinline__ d3 nim_search(const Box& b, d3 ri, d3 rj) {
    d3 r{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    if (b.kind == 2) return nim_tri(b, r);
    if (b.kind == 1) {
        const double ux = r.x * b.invL[0], uy = r.y * b.invL[1], uz = r.z * b.invL[2];
        double nx = rint(ux), ny = rint(uy), nz = rint(uz);
        const bool tie = fabs(ux - nx) == 0.5 || fabs(uy - ny) == 0.5 || fabs(uz - nz) == 0.5;
        if (__builtin_expect(tie, 0)) {
            nx = round(ux);
            ny = round(uy);
            nz = round(uz);
        }
        r.x = r.x - b.L[0] * nx;
        r.y = r.y - b.L[1] * ny;
        r.z = r.z - b.L[2] * nz;
    }
    return r;
}

template <bool FILL>
__global__ void __launch_bounds__(BLOCK) k_nbl(SysPtrs p, int32_t ncell, double near_r2) {
    __shared__ double sx[NBL_PIECE], sy[NBL_PIECE], sz[NBL_PIECE];
    __shared__ int32_t sm[NBL_PIECE], scm[NBL_PIECE], sst[NBL_PIECE], snn[NBL_PIECE];
    __shared__ int32_t run_m0[NBL_RUNS], run_off[NBL_RUNS + 1];
    // each wave's own groups (reference point, original index), read back with a
    // uniform address in the candidate loop: a broadcast through LDS instead of
    // lane reads on the vector ALU
    __shared__ double s_kx[NBL_WAVES][WAVE], s_ky[NBL_WAVES][WAVE], s_kz[NBL_WAVES][WAVE];
    __shared__ int32_t s_ck[NBL_WAVES][WAVE];
    const int32_t c = xcd_block(ncell);
    if (c < 0) return;
    const int32_t s0 = p.cell_start[c], s1 = p.cell_start[c + 1];
    if (s0 >= s1) return;     // empty cell: the whole block leaves before any barrier
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const JobDev& J = p.jobs[p.cg_job[p.s_cg_orig[s0]]];
    const Box box = J.box;
    const double c2s = J.c2s, c2l = J.c2l;
    const int32_t cg_off = J.cg_off, nsol = J.n_solute_cg, aps = J.aps;
    const int32_t nx = J.ncell[0], ny = J.ncell[1], nz = J.ncell[2];
    const int32_t local = c - J.cell_off;
    const int32_t cc[3] = {local % nx, (local / nx) % ny, local / (nx * ny)};
    // the stencil as x runs of sorted groups, in (z, y, x-segment) order
    if (wv == 0) {
        const int32_t nyy = ny >= 2 * CELL_R + 1 ? 2 * CELL_R + 1 : ny;
        const int32_t nzz = nz >= 2 * CELL_R + 1 ? 2 * CELL_R + 1 : nz;
        int32_t m0 = 0, cnt = 0;
        const int32_t r = lane, rowi = r / 2, seg = r % 2;
        if (r < NBL_RUNS && rowi < nyy * nzz) {
            const int32_t iy = rowi % nyy, iz = rowi / nyy;
            const int32_t ly = ny >= 2 * CELL_R + 1 ? (cc[1] + ny + iy - CELL_R) % ny : iy;
            const int32_t lz = nz >= 2 * CELL_R + 1 ? (cc[2] + nz + iz - CELL_R) % nz : iz;
            const int32_t row = J.cell_off + nx * (ly + ny * lz);
            int32_t a0 = 0, a1 = 0;   // cell range [a0, a1) of this segment
            if (nx < 2 * CELL_R + 1) {
                if (seg == 0) { a0 = 0; a1 = nx; }
            } else {
                const int32_t x0 = cc[0] - CELL_R, x1 = cc[0] + CELL_R;
                if (x0 < 0) {
                    if (seg == 0) { a0 = x0 + nx; a1 = nx; } else { a0 = 0; a1 = x1 + 1; }
                } else if (x1 >= nx) {
                    if (seg == 0) { a0 = x0; a1 = nx; } else { a0 = 0; a1 = x1 - nx + 1; }
                } else if (seg == 0) {
                    a0 = x0;
                    a1 = x1 + 1;
                }
            }
            if (a1 > a0) {
                m0 = p.cell_start[row + a0];
                cnt = p.cell_start[row + a1] - m0;
            }
        }
        // exclusive scan of the run lengths over the wave
        int32_t inc = cnt;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            int32_t v = __shfl_up(inc, o, WAVE);
            if (lane >= o) inc += v;
        }
        if (r < NBL_RUNS) {
            run_m0[r] = m0;
            run_off[r] = inc - cnt;
        }
        if (r == NBL_RUNS - 1) run_off[NBL_RUNS] = inc;
    }
    __syncthreads();
    const int32_t total = run_off[NBL_RUNS];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t pairs_short = 0;                        // wave-uniform: solvent pairs
    uint64_t pairs_short_l = 0, pairs_long_l = 0;    // lane sums: solute pairs
    for (int32_t rb = s0; rb < s1; rb += BLOCK) {
        // this round's groups: lane t of wave w holds group rb + w + NBL_WAVES t
        const int32_t nr = min(BLOCK, s1 - rb);
        const int32_t gi = wv + NBL_WAVES * lane;
        const bool own = gi < nr;
        const int32_t k = own ? rb + gi : rb;
        const double kx = p.s_ref[3 * k], ky = p.s_ref[3 * k + 1], kz = p.s_ref[3 * k + 2];
        const int32_t ck = p.s_cg_orig[k], nk = p.s_cg_n[k];
        const uint64_t own_mask = __ballot(own);
        const int32_t n_own = (int32_t)__popcll(own_mask);   // a prefix of the lanes
        const uint64_t solk_mask = __ballot(own && ck - cg_off < nsol);
        s_kx[wv][lane] = kx;   // read after the barrier that opens the first piece
        s_ky[wv][lane] = ky;
        s_kz[wv][lane] = kz;
        s_ck[wv][lane] = ck;
        int32_t vv_base = 0, gen_base = 0, vv_end = 0, gen_end = 0;
        if (FILL && own) {
            vv_base = (int32_t)p.vv_off[k];
            gen_base = (int32_t)p.gen_off[k];
            vv_end = vv_base + (int32_t)p.vv_cap[k];
            gen_end = gen_base + (int32_t)p.gen_cap[k];
        }
        // solvent entries: near partners (d2 < near_r2) from the slot's start
        // upwards, the others from its end downwards
        int32_t vv_nn = 0, vv_nf = 0, gen_pos = gen_base;
        for (int32_t P = 0; P < total; P += NBL_PIECE) {
            const int32_t plen = min(NBL_PIECE, total - P);
            __syncthreads();   // the previous piece is consumed
            for (int32_t q = threadIdx.x; q < plen; q += BLOCK) {
                const int32_t g = P + q;
                // run holding candidate g: the last run starting at or before g
                int32_t lo = 0, hi = NBL_RUNS - 1;
                while (lo < hi) {
                    const int32_t mid = (lo + hi + 1) >> 1;
                    if (run_off[mid] <= g) lo = mid; else hi = mid - 1;
                }
                const int32_t m = run_m0[lo] + (g - run_off[lo]);
                sx[q] = p.s_ref[3 * m];
                sy[q] = p.s_ref[3 * m + 1];
                sz[q] = p.s_ref[3 * m + 2];
                sm[q] = m;
                scm[q] = p.s_cg_orig[m];
                sst[q] = p.s_cg_start[m];
                snn[q] = p.s_cg_n[m];
            }
            __syncthreads();
            if (!own_mask) continue;   // wave-uniform; the barriers above are reached by all
            for (int32_t cb = 0; cb < plen; cb += WAVE) {
                const int32_t q = cb + lane;
                const bool valid = q < plen;
                const int32_t qq = valid ? q : 0;
                const d3 rm{sx[qq], sy[qq], sz[qq]};
                const int32_t m = sm[qq], cm = scm[qq];
                const bool sol_m = cm - cg_off < nsol;
                const uint32_t vv_ent = (uint32_t)sst[qq];
                const int32_t n_m = snn[qq];
                // the wave's own groups are lanes 0 .. n_own - 1 (gi = wv + 4 lane < nr)
                for (int32_t t = 0; t < n_own; ++t) {
                    const int32_t i = rb + wv + NBL_WAVES * t;
                    const d3 rk{s_kx[wv][t], s_ky[wv][t], s_kz[wv][t]};
                    const int32_t ck_t = s_ck[wv][t];
                    const bool sol_k = (solk_mask >> t) & 1ull;
                    const d3 r = nim_search(box, rk, rm);
                    const double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
                    const bool in = valid && m != i && !(d2 > c2l);
                    const bool lng = d2 > c2s;
                    const bool canon = ck_t < cm;
                    const bool is_vv = in && !lng && !sol_k && !sol_m;
                    const bool is_gen = in && !is_vv;
                    const uint64_t bv = __ballot(is_vv), bg = __ballot(is_gen);
                    if (bv) {
                        const bool nr = is_vv && d2 < near_r2;
                        const uint64_t bn = __ballot(nr), bf = bv & ~bn;
                        if (FILL) {
                            // group t's slot and cursors (wave-uniform), then one store:
                            // the near lanes' rank from the start, the far lanes' from the end
                            const int32_t base = bcast(vv_base, t), cap = bcast(vv_end, t) - base;
                            const int32_t nn0 = bcast(vv_nn, t), nf0 = bcast(vv_nf, t);
                            const int32_t x = nr ? nn0 + (int32_t)__popcll(bn & below)
                                                 : nf0 + (int32_t)__popcll(bf & below);
                            if (is_vv && x < cap)
                                p.vv_ent[nr ? base + x : base + cap - 1 - x] = vv_ent | (canon ? ENT_CANON : 0u);
                        }
                        if (lane == t) {
                            vv_nn += (int32_t)__popcll(bn);
                            vv_nf += (int32_t)__popcll(bf);
                        }
                        pairs_short += (uint64_t)__popcll(bv & __ballot(canon)) * (uint64_t)(aps * aps);
                    }
                    if (bg) {
                        // solute partners, and solvent pairs beyond the short cutoff
                        uint32_t flags = canon ? ENT_CANON : 0u;
                        if (!sol_k && !sol_m) flags |= ENT_VVLONG;
                        if (lng) flags |= ENT_LONG;
                        if (sol_k && sol_m) flags |= ENT_SS;
                        if (FILL) {
                            const int32_t pg = bcast(gen_pos, t) + (int32_t)__popcll(bg & below);
                            if (is_gen && pg < bcast(gen_end, t)) p.gen_ent[pg] = (uint32_t)m | flags;
                        }
                        if (lane == t) gen_pos += (int32_t)__popcll(bg);
                        if (is_gen && canon) {
                            const uint64_t np = (uint64_t)bcast(nk, t) * (uint64_t)n_m;
                            if (lng) pairs_long_l += np; else pairs_short_l += np;
                        }
                    }
                }
            }
        }
        if (own) {
            const int32_t nvv = vv_nn + vv_nf, ngen = gen_pos - gen_base;
            p.vv_cnt[k] = nvv;
            p.vv_near[k] = vv_nn;
            p.gen_cnt[k] = ngen;
            p.prev_vv[ck] = nvv;   // the next build's slot sizes (read by k_caps before its search)
            p.prev_gen[ck] = ngen;
            p.pc_short[k] = 0;
            p.pc_long[k] = 0;
            if (FILL && (nvv > vv_end - vv_base || ngen > gen_end - gen_base)) atomicOr(p.overflow, 1);
        }
        // solvent groups with general-list entries (solute partners, long pairs)
        // for the mixed solvent pass, in any order: each is evaluated on its own
        const bool gs = own && ck - cg_off >= nsol && gen_pos > gen_base;
        const uint64_t gm = __ballot(gs);
        if (gm) {
            const int32_t t0 = (int32_t)__builtin_ctzll(gm);
            int32_t base = 0;
            if (lane == t0)
                base = (int32_t)atomicAdd((unsigned long long*)p.counters + 4, (unsigned long long)__popcll(gm));
            base = bcast(base, t0);
            if (gs) p.gsol[base + (int32_t)__popcll(gm & below)] = k;
        }
    }
    // per-cell pair counts; totals come from k_totals (one atomic per block
    // there: thousands of blocks adding to the same address serialise at the L2)
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        pairs_short_l += __shfl_xor(pairs_short_l, o, WAVE);
        pairs_long_l += __shfl_xor(pairs_long_l, o, WAVE);
    }
    __shared__ uint64_t red[2 * NBL_WAVES];
    if (lane == 0) {
        red[2 * wv] = pairs_short + pairs_short_l;
        red[2 * wv + 1] = pairs_long_l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, b2 = 0;
        for (int w = 0; w < NBL_WAVES; ++w) {
            a += red[2 * w];
            b2 += red[2 * w + 1];
        }
        p.pc_short[s0] = (int64_t)a;
        p.pc_long[s0] = (int64_t)b2;
    }
}

// counters[0..3] = sums over groups of short pairs, long pairs, vv entries,
// gen entries (counters cleared before)
__global__ void __launch_bounds__(BLOCK) k_totals(SysPtrs p, int32_t ncg) {
    __shared__ double lds[17 * 4];
    double v[4] = {0.0, 0.0, 0.0, 0.0};   // exact: each sum stays below 2^53
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ncg; k += gridDim.x * blockDim.x) {
        v[0] += (double)p.pc_short[k];
        v[1] += (double)p.pc_long[k];
        v[2] += (double)p.vv_cnt[k];
        v[3] += (double)p.gen_cnt[k];
    }
    block_sum<4>(v, lds);
    if (threadIdx.x < 4) {
        unsigned long long t = (unsigned long long)lds[16 * 4 + threadIdx.x];
        if (t) atomicAdd((unsigned long long*)p.counters + threadIdx.x, t);
    }
}

// per-job counts of this build (pair interactions short, long; solvent-solvent
// list entries; job_pairs cleared before): blocks (x, j) sum slices of job j's
// sorted groups, which are contiguous from cg_off, so a step can count only the
// jobs that advance in it
constexpr int JOB_PAIR_SLICES = 16;
__global__ void __launch_bounds__(BLOCK) k_job_pairs(SysPtrs p) {
    __shared__ double lds[17 * JOB_CTL];
    const JobDev& J = p.jobs[blockIdx.y];
    const int32_t k1 = J.cg_off + J.n_cg;
    double v[JOB_CTL] = {0.0, 0.0, 0.0};   // exact: each sum stays below 2^53
    for (int32_t k = J.cg_off + (int32_t)(blockIdx.x * BLOCK + threadIdx.x); k < k1; k += JOB_PAIR_SLICES * BLOCK) {
        v[0] += (double)p.pc_short[k];
        v[1] += (double)p.pc_long[k];
        v[2] += (double)p.vv_cnt[k];
    }
    block_sum<JOB_CTL>(v, lds);
    if (threadIdx.x < JOB_CTL) {
        const unsigned long long t = (unsigned long long)lds[16 * JOB_CTL + threadIdx.x];
        if (t) atomicAdd((unsigned long long*)p.job_pairs + JOB_CTL * blockIdx.y + threadIdx.x, t);
    }
}

// capacities for a single-pass fill: the previous build's count of the same
// (original) group plus 1/8 and 16 entries
// tight (a test hook): exactly the previous counts, so that most builds refill
__global__ void k_caps(SysPtrs p, int32_t ncg, int32_t tight) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncg) return;
    int32_t c = p.s_cg_orig[k];
    int64_t v = p.prev_vv[c], g = p.prev_gen[c];
    p.vv_cap[k] = tight ? v : v + (v >> 3) + 16;
    p.gen_cap[k] = tight ? g : g + (g >> 3) + 16;
}

// capacities for an exact refill are the counts themselves
__global__ void k_caps_from_counts(SysPtrs p, int32_t ncg) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncg) return;
    p.vv_cap[k] = p.vv_cnt[k];
    p.gen_cap[k] = p.gen_cnt[k];
}

// ---------------------------------------------------------------------------
// forces
// ---------------------------------------------------------------------------

__global__ void k_gather(SysPtrs p, int32_t npad, int32_t mixed) {
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= npad) return;
    int32_t o = p.s2o[a];
    double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
    if (o >= 0) v = make_double4(p.x_cur[3 * o], p.x_cur[3 * o + 1], p.x_cur[3 * o + 2], p.q[o]);
    p.xq[a] = v;
    if (mixed) {
        // FP32 offset from the group's first atom (difference taken in FP64)
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
        if (o >= 0) {
            int32_t f = p.cg_start[p.s_cg_orig[p.s_atom_cg[a]]];
            d = make_float4((float)(v.x - p.x_cur[3 * f]), (float)(v.y - p.x_cur[3 * f + 1]),
                            (float)(v.z - p.x_cur[3 * f + 2]), (float)v.w);
        }
        p.xf[a] = d;
    }
}

__device__ __forceinline__ d3 xyz(double4 v) { return d3{v.x, v.y, v.z}; }

__device__ __forceinline__ void add_virial(double (&v)[ROW], d3 r, double fx, double fy, double fz) {
    v[2] += r.x * fx; v[3] += r.x * fy; v[4] += r.x * fz;
    v[5] += r.y * fx; v[6] += r.y * fy; v[7] += r.y * fz;
    v[8] += r.z * fx; v[9] += r.z * fy; v[10] += r.z * fz;
}

__device__ __forceinline__ void write_row(double (&v)[ROW], double* rows, int slot, int nblocks, int32_t blk) {
    __shared__ double lds[17 * ROW];
    block_sum<ROW>(v, lds);
    if (threadIdx.x < ROW)
        rows[((int64_t)slot * nblocks + blk) * ROW + threadIdx.x] = lds[16 * ROW + threadIdx.x];
}
__device__ __forceinline__ void write_row(double (&v)[ROW], double* rows, int slot, int nblocks) {
    write_row(v, rows, slot, nblocks, (int32_t)blockIdx.x);
}

// ---- mixed precision: FP32 pair arithmetic, FP64 per-lane partial sums,
// ---- int64 fixed-point per-atom totals (exact and order-independent)
constexpr double FX_SCALE = 16777216.0;          // 2^24 per kJ mol^-1 nm^-1
constexpr double FX_INV = 1.0 / 16777216.0;
constexpr double EX_SCALE = 4294967296.0;        // 2^32 per kJ mol^-1 (energies, virial)
constexpr double EX_INV = 1.0 / 4294967296.0;
constexpr int VV_GROUP = 8;                      // entries per fixed-point group

// Conversions saturate (a NaN goes to the negative cap): a diverged job's
// forces stay defined sums that k_integrate recognises (FORCE_BAD) and turns
// into a failed job.  At most 2^12 conversions add up per value, so the int64
// sums cannot wrap.
constexpr double FX_CAP = 2.0 * FORCE_LIMIT * FX_SCALE;   // 2^50: |F| of 2^26 kJ mol^-1 nm^-1 per term
constexpr double EX_CAP = 18014398509481984.0;            // 2^54: |E| of 2^22 kJ mol^-1 per term
__device__ __forceinline__ long long to_fx(double v) { return (long long)rint(fmin(fmax(v * FX_SCALE, -FX_CAP), FX_CAP)); }
__device__ __forceinline__ long long to_ex(double v) { return (long long)rint(fmin(fmax(v * EX_SCALE, -EX_CAP), EX_CAP)); }

__device__ __forceinline__ void lj_crf32(float rx, float ry, float rz, float c6, float c12, float qp, float c2,
                                         float cc, float c3, float& f, float& elj, float& ecrf) {
    float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    bool ok = !(r2 < 1e-10f);
    // hardware reciprocal square root (1 ulp) refined by one Newton step
    float y = __builtin_amdgcn_rsqf(r2);
    float inv_r = fmaf(0.5f * y, fmaf(-r2 * y, y, 1.0f), y);
    float inv_r2 = inv_r * inv_r;
    float inv_r6 = inv_r2 * inv_r2 * inv_r2;
    float e = fmaf(c12, inv_r6, -c6) * inv_r6;
    float fl = fmaf(12.0f * c12, inv_r6, -6.0f * c6) * inv_r6 * inv_r2;
    float ec = qp * (inv_r - fmaf(c2, r2, cc));
    float fc = qp * fmaf(inv_r, inv_r2, c3);
    f = ok ? fl + fc : 0.0f;
    elj = ok ? e : 0.0f;
    ecrf = ok ? ec : 0.0f;
}

// nearest-image vector between two group first atoms in FP64, returned as
// float, with the reference's image choice (ties included)
__device__ __forceinline__ void shift32(const Box& b, d3 A, d3 B, float& x, float& y, float& z) {
    double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    if (b.kind == 2) {
        const d3 r = nim_tri_fast(b, d3{dx, dy, dz});
        dx = r.x;
        dy = r.y;
        dz = r.z;
    } else if (b.kind == 1) {
        // rint, and the reference's half-away rounding (f64::round) where the
        // ratio is an exact half: a branch the wave skips unless a lane has one
        const double ux = dx * b.invL[0], uy = dy * b.invL[1], uz = dz * b.invL[2];
        double nx = rint(ux), ny = rint(uy), nz = rint(uz);
        if (__builtin_expect(fabs(ux - nx) == 0.5 || fabs(uy - ny) == 0.5 || fabs(uz - nz) == 0.5, 0)) {
            nx = round(ux);
            ny = round(uy);
            nz = round(uz);
        }
        dx = dx - b.L[0] * nx;
        dy = dy - b.L[1] * ny;
        dz = dz - b.L[2] * nz;
    }
    x = (float)dx;
    y = (float)dy;
    z = (float)dz;
}

// Solvent-solvent short pairs (process_solvent_pairs, innerloops.rs:218-309).
// One lane per (molecule, chunk): NCHM lanes share a solvent molecule, take
// every NCHM-th listed partner molecule and evaluate all APS x APS atom pairs
// with one shift per molecule pair; their partial forces are summed with a
// fixed xor butterfly, so no atomics are needed and the order is fixed.  Each
// lane works from its own side: r = (x_own + t) - x_partner with t from
// nearest_image(own first atom, partner first atom).  For the lane of the lower
// group (the "canonical" one) that is the reference's operation sequence, so
// its per-pair energies are the reference's and only it adds energies and
// virial.  LJ00: only the first-first atom pair has LJ parameters (water
// models whose hydrogens carry none); the other pairs' LJ terms are exactly
// zero in the reference and are skipped.  Energies and virial go to a per-group
// record (vv_e, 11 doubles) that k_reduce sums per job.  NCHM (lanes per
// molecule: 4, 8 or 16) is chosen per batch shape by timing (autotune).

// A job held past its stop step in the current md call (it leaves at the next
// batch change; the integrator keeps its state): its pair kernels skip it.
__device__ __forceinline__ bool held(const SysPtrs& p, int32_t job) { return *p.step >= p.jobs[job].stop; }

// Wave -> groups of the solvent pair kernels: the batch's sorted groups in
// chunks of VV_CHUNK that never straddle two jobs (vv_chunk = first group, end,
// job); a wave holds WAVE / NCHM consecutive groups of one chunk, so the job's
// constants are wave-uniform (scalar registers, fewer vector registers).
// Returns false for waves past the last chunk (the whole wave leaves).
template <int NCHM>
__device__ __forceinline__ void vv_wave_at(const SysPtrs& p, int32_t wvi, int32_t& k, bool& valid, int& chunk,
                                           int32_t& job) {
    constexpr int G = WAVE / NCHM;
    constexpr int SUB = VV_CHUNK / G;
    const int4 ch = p.vv_chunk[wvi / SUB];
    const int32_t first = __builtin_amdgcn_readfirstlane(ch.x), end = __builtin_amdgcn_readfirstlane(ch.y);
    job = __builtin_amdgcn_readfirstlane(ch.z);
    const int lane = threadIdx.x & (WAVE - 1);
    k = first + (wvi % SUB) * G + lane / NCHM;
    valid = k < end;
    chunk = lane % NCHM;
}
template <int NCHM>
__device__ __forceinline__ bool vv_wave(const SysPtrs& p, int32_t nwaves, int32_t& k, bool& valid, int& chunk,
                                        int32_t& job) {
    constexpr int G = WAVE / NCHM;
    constexpr int SUB = VV_CHUNK / G;
    const int32_t lb = xcd_block((nwaves + BLOCK / WAVE - 1) / (BLOCK / WAVE));
    if (lb < 0) return false;
    const int32_t wvi = lb * (BLOCK / WAVE) + (int32_t)(threadIdx.x / WAVE);
    if (wvi >= nwaves) return false;
    const int4 ch = p.vv_chunk[wvi / SUB];
    const int32_t first = __builtin_amdgcn_readfirstlane(ch.x), end = __builtin_amdgcn_readfirstlane(ch.y);
    job = __builtin_amdgcn_readfirstlane(ch.z);
    const int lane = threadIdx.x & (WAVE - 1);
    k = first + (wvi % SUB) * G + lane / NCHM;
    valid = k < end;
    chunk = lane % NCHM;
    return true;
}

template <int APS, bool LJ00>
__device__ __forceinline__ bool has_lj(int i, int j) {
    return !LJ00 || (i == 0 && j == 0);
}

template <int APS, bool LJ00, bool VIR, int NCHM>
__global__ void __launch_bounds__(BLOCK) k_vv(SysPtrs p, int32_t nwaves) {
    int32_t k = 0, job = 0;
    bool valid = false;
    int chunk = 0;
    if (!vv_wave<NCHM>(p, nwaves, k, valid, chunk, job)) return;
    if (held(p, job)) return;
    double f[APS][3];
    double en[11];
#pragma unroll
    for (int i = 0; i < APS; ++i) f[i][0] = f[i][1] = f[i][2] = 0.0;
#pragma unroll
    for (int c = 0; c < 11; ++c) en[c] = 0.0;
    int32_t e0 = 0, e1 = 0, sk = 0;
    if (valid) {
        e0 = (int32_t)p.vv_off[k];
        e1 = e0 + (int32_t)min(p.vv_cnt[k], p.vv_cap[k]);   // entries written (a pending refill)
        sk = p.s_cg_start[k];
    }
    if (e1 > e0) {
        const JobDev& J = p.jobs[job];
        const Box box = J.box;
        const double c2 = J.crf.crf_2cut3i, cc = J.crf.crf_cut, c3 = J.crf.crf_cut3i;
        d3 pa[APS];
#pragma unroll
        for (int i = 0; i < APS; ++i) pa[i] = xyz(p.xq[sk + i]);
        for (int32_t e = e0 + chunk; e < e1; e += NCHM) {
            uint32_t ent = p.vv_ent[e];
            int32_t sm = (int32_t)(ent & ENT_IDX);
            const bool canon = (ent & ENT_CANON) != 0;
            d3 pb[APS];
#pragma unroll
            for (int j = 0; j < APS; ++j) pb[j] = xyz(p.xq[sm + j]);
            d3 rf = nim(box, pa[0], pb[0]);
            double tx = rf.x - pa[0].x + pb[0].x, ty = rf.y - pa[0].y + pb[0].y, tz = rf.z - pa[0].z + pb[0].z;
#pragma unroll
            for (int i = 0; i < APS; ++i) {
                double xi = pa[i].x + tx, yi = pa[i].y + ty, zi = pa[i].z + tz;
#pragma unroll
                for (int j = 0; j < APS; ++j) {
                    double rx = xi - pb[j].x, ry = yi - pb[j].y, rz = zi - pb[j].z;
                    // lj_crf_interaction (innerloops.rs:22-55), branch-free guard
                    double r2 = rx * rx + ry * ry + rz * rz;
                    bool ok = !(r2 < 1e-10);
                    double inv_r2 = 1.0 / r2;
                    double inv_r = sqrt(inv_r2);
                    double qq = J.sv_qq[i * 4 + j];
                    double ecrf = qq * (inv_r - c2 * r2 - cc);
                    double fm = qq * (inv_r * inv_r2 + c3);
                    double elj = 0.0;
                    if (has_lj<APS, LJ00>(i, j)) {
                        double c6 = J.sv_c6[i * 4 + j], c12 = J.sv_c12[i * 4 + j];
                        double inv_r6 = inv_r2 * inv_r2 * inv_r2;
                        elj = (c12 * inv_r6 - c6) * inv_r6;
                        fm = (12.0 * c12 * inv_r6 - 6.0 * c6) * inv_r6 * inv_r2 + fm;
                    }
                    fm = ok ? fm : 0.0;
                    double gx = rx * fm, gy = ry * fm, gz = rz * fm;
                    f[i][0] += gx;
                    f[i][1] += gy;
                    f[i][2] += gz;
                    if (canon) {
                        en[0] += ok ? elj : 0.0;
                        en[1] += ok ? ecrf : 0.0;
                        if (VIR) {
                            en[2] += rx * gx; en[3] += rx * gy; en[4] += rx * gz;
                            en[5] += ry * gx; en[6] += ry * gy; en[7] += ry * gz;
                            en[8] += rz * gx; en[9] += rz * gy; en[10] += rz * gz;
                        }
                    }
                }
            }
        }
    }
    // fixed-order butterfly over the NCHM lanes of one molecule
#pragma unroll
    for (int m = 1; m < NCHM; m <<= 1) {
#pragma unroll
        for (int i = 0; i < APS; ++i) {
            f[i][0] += __shfl_xor(f[i][0], m, WAVE);
            f[i][1] += __shfl_xor(f[i][1], m, WAVE);
            f[i][2] += __shfl_xor(f[i][2], m, WAVE);
        }
#pragma unroll
        for (int c = 0; c < (VIR ? 11 : 2); ++c) en[c] += __shfl_xor(en[c], m, WAVE);
    }
    if (chunk == 0 && valid) {
        int32_t nk = p.s_cg_n[k];
        if (e1 > e0) {
#pragma unroll
            for (int i = 0; i < APS; ++i) {
                p.f_s[3 * (sk + i)] = f[i][0];
                p.f_s[3 * (sk + i) + 1] = f[i][1];
                p.f_s[3 * (sk + i) + 2] = f[i][2];
            }
        } else {
            for (int32_t i = 0; i < 3 * nk; ++i) p.f_s[3 * sk + i] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < 11; ++c) p.vv_e[11 * (int64_t)k + c] = en[c];
    }
}

// Mixed-precision counterpart: FP32 pair arithmetic on offsets from the
// molecules' first atoms (their nearest-image vector taken in FP64), FP64 per-
// group sums, int64 fixed-point per-atom result.
//
// NEUT (every solvent molecule has zero net charge, as SPC water): the reaction-
// field polynomial is summed per molecule pair instead of per atom pair.  With
// r_ij = u_i - w_j and sum_i q_i = sum_j q_j = 0,
//   sum_ij q_i q_j (-crf_2cut3i r_ij^2 - crf_cut) =  2 crf_2cut3i  mu_A . mu_B
//   sum_j  q_i q_j  crf_cut3i r_ij               = -q_i crf_cut3i  mu_B
//   sum_ij q_i q_j  crf_cut3i r_ij (x) r_ij       = -crf_cut3i (mu_A (x) mu_B + mu_B (x) mu_A)
// where mu = sum q x is the molecule's dipole (origin independent for a neutral
// molecule), so per atom pair only the Coulomb 1/r and LJ terms remain.  This
// is the same sum as the reference's n x n loop (innerloops.rs:256-299) in
// exact arithmetic, and in floating point it drops the cancelling constant
// terms.  Atom pairs closer than 1e-5 nm (never two molecules) keep zero 1/r
// terms but, unlike the reference, their RF polynomial is not dropped.
//
// Near partners: the search writes the partners whose first atoms are closer
// than VV_NEAR_R at the slot's start (upwards) and the others at its end
// (downwards); this kernel evaluates the latter, k_vv_nr the near ones in FP64.
template <int APS, bool LJ00, bool VIR, int NCHM, bool NEUT>
__device__ __forceinline__ void vv_mx_wave(const SysPtrs& p, int32_t k, bool valid, int chunk, int32_t job) {
    if (held(p, job)) return;
    long long fi[APS][3];
    long long ei[11];
#pragma unroll
    for (int i = 0; i < APS; ++i) fi[i][0] = fi[i][1] = fi[i][2] = 0;
#pragma unroll
    for (int c = 0; c < 11; ++c) ei[c] = 0;
    // the far partners: the last min(count - near, cap) entries of the slot (a
    // pending refill may have dropped some)
    int32_t e1 = 0, ef = 0, sk = 0;
    int64_t cnt = 0;
    if (valid) {
        const int64_t cap = p.vv_cap[k];
        cnt = p.vv_cnt[k];
        e1 = (int32_t)(p.vv_off[k] + cap);
        ef = e1 - (int32_t)min(cnt - (int64_t)p.vv_near[k], cap);
        sk = p.s_cg_start[k];
    }
    if (e1 > ef) {
        const JobDev& J = p.jobs[job];
        const Box box = J.box;
        const float c2 = (float)J.crf.crf_2cut3i, cc = (float)J.crf.crf_cut, c3 = (float)J.crf.crf_cut3i;
        float qq[APS][APS];
#pragma unroll
        for (int i = 0; i < APS; ++i)
#pragma unroll
            for (int j = 0; j < APS; ++j) qq[i][j] = (float)J.sv_qq[i * 4 + j];
        const float c6_00 = (float)J.sv_c6[0], c12_00 = (float)J.sv_c12[0];
        float4 da[APS];
#pragma unroll
        for (int i = 0; i < APS; ++i) da[i] = p.xf[sk + i];
        // NEUT: own dipole and the per-atom RF force factor -q_i fpepsi crf_cut3i
        double mua[3] = {0.0, 0.0, 0.0};
        double kf[APS];
        if (NEUT) {
#pragma unroll
            for (int i = 1; i < APS; ++i) {
                mua[0] += (double)da[i].w * (double)da[i].x;
                mua[1] += (double)da[i].w * (double)da[i].y;
                mua[2] += (double)da[i].w * (double)da[i].z;
            }
#pragma unroll
            for (int i = 0; i < APS; ++i) kf[i] = -(double)da[i].w * J.fpepsi * J.crf.crf_cut3i;
        }
        d3 pk0 = xyz(p.xq[sk]);
        const int32_t n = e1 - ef;
        // Groups of VV_GROUP consecutive entries are summed (forces in FP32: every
        // partner here is beyond the near radius, so the terms are moderate) and
        // then added as fixed point: the result does not depend on how groups
        // are spread over the lanes, so every lane count gives the same bits.
        for (int32_t g0 = chunk * VV_GROUP; g0 < n; g0 += NCHM * VV_GROUP) {
            float fs[APS][3];
            double en[11];
#pragma unroll
            for (int i = 0; i < APS; ++i) fs[i][0] = fs[i][1] = fs[i][2] = 0.f;
#pragma unroll
            for (int c = 0; c < 11; ++c) en[c] = 0.0;
            float msx = 0.f, msy = 0.f, msz = 0.f;   // NEUT: sum of partner dipoles
            float mcx = 0.f, mcy = 0.f, mcz = 0.f;   //       over canonical entries
            float els = 0.f, ecs = 0.f;              // canonical entries' energies
            const int32_t g1 = min(g0 + VV_GROUP, n);
            for (int32_t e = ef + g0; e < ef + g1; ++e) {
                uint32_t ent = p.vv_ent[e];
                int32_t sm = (int32_t)(ent & ENT_IDX);
                const float canon = (ent & ENT_CANON) ? 1.0f : 0.0f;
                float sx, sy, sz;
                shift32(box, pk0, xyz(p.xq[sm]), sx, sy, sz);
                // a first atom's offset from itself is exactly zero (every writer
                // of xf subtracts the same FP64 value), so atom 0's slot is not
                // read and its zero offsets are left out of the arithmetic below
                float4 db[APS];
                db[0] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int j = 1; j < APS; ++j) db[j] = p.xf[sm + j];
                if (NEUT) {
                    float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
                    for (int j = 1; j < APS; ++j) {
                        ux = fmaf(db[j].w, db[j].x, ux);
                        uy = fmaf(db[j].w, db[j].y, uy);
                        uz = fmaf(db[j].w, db[j].z, uz);
                    }
                    msx += ux;
                    msy += uy;
                    msz += uz;
                    mcx = fmaf(canon, ux, mcx);
                    mcy = fmaf(canon, uy, mcy);
                    mcz = fmaf(canon, uz, mcz);
                }
                float el = 0.f, ec = 0.f;
                float w[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int i = 0; i < APS; ++i) {
                    const float bx = i ? sx + da[i].x : sx, by = i ? sy + da[i].y : sy, bz = i ? sz + da[i].z : sz;
                    float gxs = fs[i][0], gys = fs[i][1], gzs = fs[i][2];
#pragma unroll
                    for (int j = 0; j < APS; ++j) {
                        const float rx = j ? bx - db[j].x : bx, ry = j ? by - db[j].y : by, rz = j ? bz - db[j].z : bz;
                        float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
                        bool ok = !(r2 < 1e-10f);
                        // the hardware reciprocal square root: within 0.87 ulp over
                        // every float in [0.125, 4) nm^2, 89 % correctly rounded
                        // (measured on gfx950), as accurate here as a Newton step
                        float inv_r = __builtin_amdgcn_rsqf(r2);
                        float q = qq[i][j];
                        float fm;
                        if (NEUT) {
                            inv_r = ok ? inv_r : 0.0f;
                            float inv_r2 = inv_r * inv_r;
                            ec = fmaf(q, inv_r, ec);
                            fm = q * (inv_r * inv_r2);
                            if (has_lj<APS, LJ00>(i, j)) {
                                float c6 = LJ00 ? c6_00 : (float)J.sv_c6[i * 4 + j];
                                float c12 = LJ00 ? c12_00 : (float)J.sv_c12[i * 4 + j];
                                float inv_r6 = inv_r2 * inv_r2 * inv_r2;
                                el = fmaf(fmaf(c12, inv_r6, -c6), inv_r6, el);
                                fm = fmaf(fmaf(12.0f * c12, inv_r6, -6.0f * c6) * inv_r6, inv_r2, fm);
                            }
                        } else {
                            float inv_r2 = inv_r * inv_r;
                            float ecr = q * (inv_r - fmaf(c2, r2, cc));
                            fm = q * fmaf(inv_r, inv_r2, c3);
                            float elj = 0.f;
                            if (has_lj<APS, LJ00>(i, j)) {
                                float c6 = LJ00 ? c6_00 : (float)J.sv_c6[i * 4 + j];
                                float c12 = LJ00 ? c12_00 : (float)J.sv_c12[i * 4 + j];
                                float inv_r6 = inv_r2 * inv_r2 * inv_r2;
                                elj = fmaf(c12, inv_r6, -c6) * inv_r6;
                                fm = fmaf(fmaf(12.0f * c12, inv_r6, -6.0f * c6) * inv_r6, inv_r2, fm);
                            }
                            fm = ok ? fm : 0.0f;
                            el += ok ? elj : 0.0f;
                            ec += ok ? ecr : 0.0f;
                        }
                        if (VIR) {
                            float gx = rx * fm, gy = ry * fm, gz = rz * fm;
                            gxs += gx;
                            gys += gy;
                            gzs += gz;
                            w[0] = fmaf(rx, gx, w[0]); w[1] = fmaf(rx, gy, w[1]); w[2] = fmaf(rx, gz, w[2]);
                            w[3] = fmaf(ry, gx, w[3]); w[4] = fmaf(ry, gy, w[4]); w[5] = fmaf(ry, gz, w[5]);
                            w[6] = fmaf(rz, gx, w[6]); w[7] = fmaf(rz, gy, w[7]); w[8] = fmaf(rz, gz, w[8]);
                        } else {
                            gxs = fmaf(rx, fm, gxs);
                            gys = fmaf(ry, fm, gys);
                            gzs = fmaf(rz, fm, gzs);
                        }
                    }
                    fs[i][0] = gxs;
                    fs[i][1] = gys;
                    fs[i][2] = gzs;
                }
                els = fmaf(canon, el, els);
                ecs = fmaf(canon, ec, ecs);
                if (VIR) {
#pragma unroll
                    for (int c = 0; c < 9; ++c) en[2 + c] += (double)(canon * w[c]);
                }
            }
            en[0] += (double)els;
            en[1] += (double)ecs;
            double f[APS][3];
#pragma unroll
            for (int i = 0; i < APS; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) f[i][c] = (double)fs[i][c];
            if (NEUT) {
                const double ms[3] = {(double)msx, (double)msy, (double)msz};
                const double mc[3] = {(double)mcx, (double)mcy, (double)mcz};
#pragma unroll
                for (int i = 0; i < APS; ++i)
#pragma unroll
                    for (int c = 0; c < 3; ++c) f[i][c] += kf[i] * ms[c];
                const double fe = J.fpepsi;
                en[1] += 2.0 * J.crf.crf_2cut3i * fe * (mua[0] * mc[0] + mua[1] * mc[1] + mua[2] * mc[2]);
                if (VIR) {
                    const double kv = -J.crf.crf_cut3i * fe;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) en[2 + 3 * a + b] += kv * (mua[a] * mc[b] + mc[a] * mua[b]);
                }
            }
#pragma unroll
            for (int i = 0; i < APS; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) fi[i][c] += to_fx(f[i][c]);
#pragma unroll
            for (int c = 0; c < (VIR ? 11 : 2); ++c) ei[c] += to_ex(en[c]);
        }
    }
#pragma unroll
    for (int m = 1; m < NCHM; m <<= 1) {
#pragma unroll
        for (int i = 0; i < APS; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) fi[i][c] += __shfl_xor(fi[i][c], m, WAVE);
#pragma unroll
        for (int c = 0; c < (VIR ? 11 : 2); ++c) ei[c] += __shfl_xor(ei[c], m, WAVE);
    }
    if (chunk == 0 && valid) {
        int32_t nk = p.s_cg_n[k];
        if (cnt > 0) {
#pragma unroll
            for (int i = 0; i < APS; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) p.fi_s[3 * (sk + i) + c] = fi[i][c];
        } else if (p.s_cg_orig[k] - p.jobs[job].cg_off >= p.jobs[job].n_solute_cg) {
            // a solvent group without partners; solute atoms' forces are assigned
            // by k_gen_sol (which may run concurrently with this kernel)
            for (int32_t i = 0; i < 3 * nk; ++i) p.fi_s[3 * sk + i] = 0;
        }
#pragma unroll
        for (int c = 0; c < 11; ++c) p.vv_e[11 * (int64_t)k + c] = (double)ei[c] * EX_INV;
    }
}

// four waves per SIMD (at most 128 vector registers, no spills)
template <int APS, bool LJ00, bool VIR, int NCHM, bool NEUT>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4))) k_vv_mx(SysPtrs p, int32_t nwaves) {
    int32_t k = 0, job = 0;
    bool valid = false;
    int chunk = 0;
    if (!vv_wave<NCHM>(p, nwaves, k, valid, chunk, job)) return;
    vv_mx_wave<APS, LJ00, VIR, NCHM, NEUT>(p, k, valid, chunk, job);
}

// The same waves in a persistent grid (as many blocks as stay resident): the
// logical waves are split over the eight XCDs as k_vv_mx's blocks are (the
// same slab of sorted groups per L2), and each resident wave takes the next
// logical wave of its XCD from a counter until none is left, so a wave that
// finishes early takes more work instead of idling until its block ends.
// Every logical wave writes only its own rows, so the order of taking changes
// no result.  sched[0..7]: the XCDs' next logical wave, sched[8]: resident
// waves finished; the last one out zeroes them for the next launch.  Every
// wave leaves: the counters only grow, and a wave leaves at the first index
// past its XCD's share.  Three waves per SIMD: with the loop around the body
// four (128 registers) spill to scratch.
template <int APS, bool LJ00, bool VIR, int NCHM, bool NEUT>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_vv_mx_persist(SysPtrs p, int32_t nwaves, int32_t per_xcd) {
    const int32_t x = (int32_t)blockIdx.x % NXCD;
    const int32_t lo = x * per_xcd, hi = min(nwaves, lo + per_xcd);
    unsigned* sched = p.vv_sched;
    const bool lead = (threadIdx.x & (WAVE - 1)) == 0;
    for (;;) {
        int32_t i = 0;
        if (lead) i = (int32_t)atomicAdd(sched + x, 1u);
        i = __builtin_amdgcn_readfirstlane(i);   // lane 0's: every lane is active here
        if (lo + i >= hi) break;
        int32_t k = 0, job = 0;
        bool valid = false;
        int chunk = 0;
        vv_wave_at<NCHM>(p, lo + i, k, valid, chunk, job);
        vv_mx_wave<APS, LJ00, VIR, NCHM, NEUT>(p, k, valid, chunk, job);
    }
    if (lead) {
        __threadfence();
        const unsigned waves = gridDim.x * (BLOCK / WAVE);
        if (atomicAdd(sched + NXCD, 1u) == waves - 1u)
            for (int c = 0; c <= NXCD; ++c) atomicExch(sched + c, 0u);
    }
}

// Near solvent partners of the mixed path (the neighbours within VV_NEAR_R,
// listed at the start of each slot by the search) in FP64 with k_vv's per-atom-pair form,
// with outputs of its own: FP32 rounding of the large, mutually
// cancelling Coulomb terms of close molecule pairs is what limits the mixed
// force error, and these pairs are ~8 % of a 1.4 nm list.  A fixed 4 lanes per
// molecule; each lane's FP64 sums become fixed point, and the group's first lane
// writes them to fi_near / vv_en, which the integrator and the energy reduction
// add to k_vv_mx's outputs.  A kernel of its own (inside k_vv_mx the FP64
// registers spill), on the side stream beside k_vv_mx in md steps.
template <int APS, bool LJ00, bool VIR, int NCHM>
__global__ void __launch_bounds__(BLOCK) k_vv_nr(SysPtrs p, int32_t nwaves) {
    int32_t k = 0, job = 0;
    bool valid = false;
    int chunk = 0;
    if (!vv_wave<NCHM>(p, nwaves, k, valid, chunk, job)) return;
    if (held(p, job)) return;
    double f[APS][3];
    double en[11];
#pragma unroll
    for (int i = 0; i < APS; ++i) f[i][0] = f[i][1] = f[i][2] = 0.0;
#pragma unroll
    for (int c = 0; c < 11; ++c) en[c] = 0.0;
    int32_t e0 = 0, nn = 0, sk = 0;
    if (valid) {
        e0 = (int32_t)p.vv_off[k];
        nn = (int32_t)min((int64_t)p.vv_near[k], p.vv_cap[k]);   // the slot's first entries
        sk = p.s_cg_start[k];
    }
    if (nn > 0) {
        const JobDev& J = p.jobs[job];
        const Box box = J.box;
        const double c2 = J.crf.crf_2cut3i, cc = J.crf.crf_cut, c3 = J.crf.crf_cut3i;
        d3 pa[APS];
#pragma unroll
        for (int i = 0; i < APS; ++i) pa[i] = xyz(p.xq[sk + i]);
        for (int32_t e = e0 + chunk; e < e0 + nn; e += NCHM) {
            const uint32_t ent = p.vv_ent[e];
            const int32_t sm = (int32_t)(ent & ENT_IDX);
            const double canon = (ent & ENT_CANON) ? 1.0 : 0.0;
            d3 pb[APS];
#pragma unroll
            for (int j = 0; j < APS; ++j) pb[j] = xyz(p.xq[sm + j]);
            const d3 rf = nim(box, pa[0], pb[0]);
            const double tx = rf.x - pa[0].x + pb[0].x, ty = rf.y - pa[0].y + pb[0].y, tz = rf.z - pa[0].z + pb[0].z;
#pragma unroll
            for (int i = 0; i < APS; ++i) {
                const double xi = pa[i].x + tx, yi = pa[i].y + ty, zi = pa[i].z + tz;
#pragma unroll
                for (int j = 0; j < APS; ++j) {
                    const double rx = xi - pb[j].x, ry = yi - pb[j].y, rz = zi - pb[j].z;
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    const bool ok = !(r2 < 1e-10);
                    // 1/r: the FP32 hardware estimate (1 ulp) and one Newton step in
                    // FP64 (relative error ~1e-14)
                    const double y0 = (double)__builtin_amdgcn_rsqf((float)r2);
                    const double inv_r = y0 * (1.5 - (0.5 * r2 * y0) * y0);
                    const double inv_r2 = inv_r * inv_r;
                    const double qq = J.sv_qq[i * 4 + j];
                    const double ecrf = qq * (inv_r - c2 * r2 - cc);
                    double fm = qq * (inv_r * inv_r2 + c3);
                    double elj = 0.0;
                    if (has_lj<APS, LJ00>(i, j)) {
                        const double c6 = J.sv_c6[i * 4 + j], c12 = J.sv_c12[i * 4 + j];
                        const double inv_r6 = inv_r2 * inv_r2 * inv_r2;
                        elj = (c12 * inv_r6 - c6) * inv_r6;
                        fm = (12.0 * c12 * inv_r6 - 6.0 * c6) * inv_r6 * inv_r2 + fm;
                    }
                    fm = ok ? fm : 0.0;
                    const double gx = rx * fm, gy = ry * fm, gz = rz * fm;
                    f[i][0] += gx;
                    f[i][1] += gy;
                    f[i][2] += gz;
                    en[0] += ok ? canon * elj : 0.0;
                    en[1] += ok ? canon * ecrf : 0.0;
                    if (VIR) {
                        const double cx = canon * rx, cy = canon * ry, cz = canon * rz;
                        en[2] += cx * gx; en[3] += cx * gy; en[4] += cx * gz;
                        en[5] += cy * gx; en[6] += cy * gy; en[7] += cy * gz;
                        en[8] += cz * gx; en[9] += cz * gy; en[10] += cz * gz;
                    }
                }
            }
        }
    }
    // each lane's FP64 sums as fixed point (the lane count is fixed at NCHM, so
    // a list gives the same bits whatever batch it is in), then the lanes' sum
    long long fi[APS][3];
    long long ei[11];
#pragma unroll
    for (int i = 0; i < APS; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) fi[i][c] = to_fx(f[i][c]);
#pragma unroll
    for (int c = 0; c < 11; ++c) ei[c] = to_ex(en[c]);
#pragma unroll
    for (int m = 1; m < NCHM; m <<= 1) {
#pragma unroll
        for (int i = 0; i < APS; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) fi[i][c] += __shfl_xor(fi[i][c], m, WAVE);
#pragma unroll
        for (int c = 0; c < (VIR ? 11 : 2); ++c) ei[c] += __shfl_xor(ei[c], m, WAVE);
    }
    // assigned for every group (zeros without near partners), so that this kernel
    // shares no output with k_vv_mx and may run beside it
    if (chunk == 0 && valid) {
        const int32_t nk = p.s_cg_n[k];
        if (nn > 0) {
#pragma unroll
            for (int i = 0; i < APS; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) p.fi_near[3 * (sk + i) + c] = fi[i][c];
        } else {
            for (int32_t i = 0; i < 3 * nk; ++i) p.fi_near[3 * sk + i] = 0;
        }
#pragma unroll
        for (int c = 0; c < 11; ++c) p.vv_en[11 * (int64_t)k + c] = (VIR || c < 2) ? (double)ei[c] * EX_INV : 0.0;
    }
}

// Topology::is_excluded_or_14 (topology.rs:853-865) on original atom indices
__device__ __forceinline__ bool bsearch32(const int32_t* a, int32_t n, int32_t key) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == key;
}
__device__ __forceinline__ bool excluded_or_14(const SysPtrs& p, int32_t i, int32_t j) {
    if (bsearch32(p.excl + p.excl_off[i], p.excl_off[i + 1] - p.excl_off[i], j)) return true;
    if (bsearch32(p.excl + p.excl_off[j], p.excl_off[j + 1] - p.excl_off[j], i)) return true;
    for (int32_t k = p.o14_off[i]; k < p.o14_off[i + 1]; ++k) if (p.o14[k] == j) return true;
    for (int32_t k = p.o14_off[j]; k < p.o14_off[j + 1]; ++k) if (p.o14[k] == i) return true;
    return false;
}

// ---------------------------------------------------------------------------
// perturbation (free energy).  The reference evaluates every pair in state A
// and then adds (perturbed - state A) for the listed pairs whose first atom is
// perturbed, for the excluded and 1-4 pairs with a perturbed atom, the RF self
// terms and the PERTATOMPAIR entries (forcefield.rs:764-951).  Here each lane
// adds those deltas for its own atom in the pass that evaluates the regular
// pair (k_gen / k_gen_mx / k_excl14), so no atomics are needed and the result
// does not depend on scheduling; the pair's first lane adds the energies and
// dH/dlambda (row column 11).
// ---------------------------------------------------------------------------
struct PertPair {
    const double* la;   // state-A LJ entry (c6, c12, cs6, cs12)
    const double* lb;   // state-B
    double a_q, b_q, alj, acrf;
};

// dual-topology parameters of the pair (i first) (perturbed.rs:335-355, 531-562)
__device__ __forceinline__ PertPair pert_params(const SysPtrs& p, const JobDev& J, int32_t i, int32_t j) {
    bool pi = p.pflag[i] != 0, pj = p.pflag[j] != 0;
    int32_t a_iac_i = pi ? p.pa_iac[i] : p.iac[i], b_iac_i = pi ? p.pb_iac[i] : p.iac[i];
    int32_t a_iac_j = pj ? p.pa_iac[j] : p.iac[j], b_iac_j = pj ? p.pb_iac[j] : p.iac[j];
    PertPair q;
    q.la = p.lj + J.lj_off + 4 * ((int64_t)a_iac_i * J.n_types + a_iac_j);
    q.lb = p.lj + J.lj_off + 4 * ((int64_t)b_iac_i * J.n_types + b_iac_j);
    q.a_q = (pi ? p.pa_q[i] : p.q[i]) * (pj ? p.pa_q[j] : p.q[j]);
    q.b_q = (pi ? p.pb_q[i] : p.q[i]) * (pj ? p.pb_q[j] : p.q[j]);
    if (pi && pj) {
        q.alj = (p.palj[i] + p.palj[j]) * 0.5;
        q.acrf = (p.pacrf[i] + p.pacrf[j]) * 0.5;
    } else {
        int32_t w = pi ? i : j;
        q.alj = p.palj[w];
        q.acrf = p.pacrf[w];
    }
    return q;
}

__device__ __forceinline__ void own_side(bool first, d3 r, double d, double& fx, double& fy, double& fz) {
    double gx = r.x * d, gy = r.y * d, gz = r.z * d;
    if (first) {
        fx += gx; fy += gy; fz += gz;
    } else {
        fx -= gx; fy -= gy; fz -= gz;
    }
}

// perturbed_pairlist_correction body for one listed pair, i first and perturbed
// (perturbed.rs:327-384)
__device__ __forceinline__ void pert_pair_delta(const SysPtrs& p, const JobDev& J, int32_t i, int32_t j, bool first,
                                                double (&v)[ROW], double& fx, double& fy, double& fz) {
    PertPair q = pert_params(p, J, i, j);
    d3 r = nim(J.box, ld3(p.x_cur, i), ld3(p.x_cur, j));
    double fa, elja, ecrfa, o[5];
    lj_crf64(r.x, r.y, r.z, q.la[0], q.la[1], q.a_q * J.fpepsi, J.crf, fa, elja, ecrfa);
    pert_lj_crf64(r.x, r.y, r.z, q.la[0], q.la[1], q.lb[0], q.lb[1], q.a_q, q.b_q, q.alj, q.acrf, J.lp, J.crf,
                  J.fpepsi, o);
    own_side(first, r, o[0] - fa, fx, fy, fz);
    if (first) {
        v[0] += o[1] - elja;
        v[1] += o[2] - ecrfa;
        v[11] += o[3] + o[4];
    }
}

// A listed short pair whose first atom i is perturbed, evaluated directly: the
// soft-core result of perturbed_lj_crf_interaction (perturbed.rs:124-266) is the
// pair's force and energy.  The reference takes the state-A pair in its regular
// loop and adds (perturbed - state A) at the nearest-image vector
// (perturbed.rs:356-380); both are the same sum, but where a soft-core atom
// overlaps another the unsoftened state-A r^-12 terms (1e15 kJ mol^-1 nm^-1 at
// 0.03 nm) leave rounding noise of thousands after the subtraction, because the
// regular loop forms r from the charge groups' shift: here nothing unsoftened
// is formed.  The virial keeps the reference's state-A pair term (the
// correction does not touch the virial).
__device__ __forceinline__ void pert_pair_direct(const SysPtrs& p, const JobDev& J, int32_t i, int32_t j, bool first,
                                                 bool vir, double (&v)[ROW], double& fx, double& fy, double& fz) {
    PertPair q = pert_params(p, J, i, j);
    d3 r = nim(J.box, ld3(p.x_cur, i), ld3(p.x_cur, j));
    double o[5];
    pert_lj_crf64(r.x, r.y, r.z, q.la[0], q.la[1], q.lb[0], q.lb[1], q.a_q, q.b_q, q.alj, q.acrf, J.lp, J.crf,
                  J.fpepsi, o);
    own_side(first, r, o[0], fx, fy, fz);
    if (first) {
        v[0] += o[1];
        v[1] += o[2];
        v[11] += o[3] + o[4];
        if (vir) {
            double fa, elja, ecrfa;
            lj_crf64(r.x, r.y, r.z, q.la[0], q.la[1], q.a_q * J.fpepsi, J.crf, fa, elja, ecrfa);
            add_virial(v, r, r.x * fa, r.y * fa, r.z * fa);
        }
    }
}

// One perturbed atom's dual-topology parameters and position, loaded once where
// it is the first atom of many pairs (a solute atom's lanes), with the soft-core
// RF constants of its pairs with unperturbed partners (alpha = its own).
struct PertAtom {
    int32_t a_iac, b_iac;
    double a_q, b_q, alj, acrf;
    d3 x;
    SoftRF rf;
};
__device__ __forceinline__ PertAtom pert_atom(const SysPtrs& p, const JobDev& J, int32_t i) {
    PertAtom a;
    a.a_iac = p.pa_iac[i];
    a.b_iac = p.pb_iac[i];
    a.a_q = p.pa_q[i];
    a.b_q = p.pb_q[i];
    a.alj = p.palj[i];
    a.acrf = p.pacrf[i];
    a.x = ld3(p.x_cur, i);
    a.rf = soft_rf(a.acrf, J.lp, J.crf);
    return a;
}

// pert_pair_direct with the first atom i given as a PertAtom (same operations)
__device__ __forceinline__ void pert_pair_direct_i(const SysPtrs& p, const JobDev& J, const PertAtom& A, int32_t j,
                                                   bool vir, double (&v)[ROW], double& fx, double& fy, double& fz) {
    const bool pj = p.pflag[j] != 0;
    const int32_t a_iac_j = pj ? p.pa_iac[j] : p.iac[j], b_iac_j = pj ? p.pb_iac[j] : p.iac[j];
    const double* la = p.lj + J.lj_off + 4 * ((int64_t)A.a_iac * J.n_types + a_iac_j);
    const double* lb = p.lj + J.lj_off + 4 * ((int64_t)A.b_iac * J.n_types + b_iac_j);
    const double a_q = A.a_q * (pj ? p.pa_q[j] : p.q[j]);
    const double b_q = A.b_q * (pj ? p.pb_q[j] : p.q[j]);
    double alj = A.alj, acrf = A.acrf;
    SoftRF rf = A.rf;
    if (pj) {
        alj = (A.alj + p.palj[j]) * 0.5;
        acrf = (A.acrf + p.pacrf[j]) * 0.5;
        rf = soft_rf(acrf, J.lp, J.crf);
    }
    d3 r = nim(J.box, A.x, ld3(p.x_cur, j));
    double o[5];
    pert_lj_crf64_rf(r.x, r.y, r.z, la[0], la[1], lb[0], lb[1], a_q, b_q, alj, acrf, J.lp, J.crf, rf, J.fpepsi, o);
    own_side(true, r, o[0], fx, fy, fz);
    v[0] += o[1];
    v[1] += o[2];
    v[11] += o[3] + o[4];
    if (vir) {
        double fa, elja, ecrfa;
        lj_crf64(r.x, r.y, r.z, la[0], la[1], a_q * J.fpepsi, J.crf, fa, elja, ecrfa);
        add_virial(v, r, r.x * fa, r.y * fa, r.z * fa);
    }
}

// pert_pair_direct_i with an unperturbed partner read from the sorted copies
// (xq = position and charge, s_type = type; the same values as the original
// arrays): one dependent load less per pair
__device__ __forceinline__ void pert_pair_direct_s(const SysPtrs& p, const JobDev& J, const PertAtom& A, double4 xb,
                                                   int32_t tb, bool vir, double (&v)[ROW], double& fx, double& fy,
                                                   double& fz) {
    const double* la = p.lj + J.lj_off + 4 * ((int64_t)A.a_iac * J.n_types + tb);
    const double* lb = p.lj + J.lj_off + 4 * ((int64_t)A.b_iac * J.n_types + tb);
    const double a_q = A.a_q * xb.w;
    const double b_q = A.b_q * xb.w;
    d3 r = nim(J.box, A.x, xyz(xb));
    double o[5];
    pert_lj_crf64_rf(r.x, r.y, r.z, la[0], la[1], lb[0], lb[1], a_q, b_q, A.alj, A.acrf, J.lp, J.crf, A.rf, J.fpepsi,
                     o);
    own_side(true, r, o[0], fx, fy, fz);
    v[0] += o[1];
    v[1] += o[2];
    v[11] += o[3] + o[4];
    if (vir) {
        double fa, elja, ecrfa;
        lj_crf64(r.x, r.y, r.z, la[0], la[1], a_q * J.fpepsi, J.crf, fa, elja, ecrfa);
        add_virial(v, r, r.x * fa, r.y * fa, r.z * fa);
    }
}

// The unperturbed second atom j's share of a listed short pair whose first atom
// i is perturbed, in FP32 (mixed path): force only; the first atom's lanes add
// the energies and their own share in FP64.  i's parameters come from its FP32
// record prec[2w..2w+1], its position from the sorted copy; j's type and charge
// are the lane's own.
__device__ __forceinline__ void pert_share32(const SysPtrs& p, const JobDev& J, int64_t w, d3 xi, d3 xj, int32_t tj,
                                             float qj, double& fx, double& fy, double& fz) {
    const float4 r0 = p.prec[2 * w], r1 = p.prec[2 * w + 1];
    const int32_t a_iac = __float_as_int(r1.z), b_iac = __float_as_int(r1.w);
    const double* la = p.lj + J.lj_off + 4 * ((int64_t)a_iac * J.n_types + tj);
    const double* lb = p.lj + J.lj_off + 4 * ((int64_t)b_iac * J.n_types + tj);
    d3 r = nim(J.box, xi, xj);
    const float rx = (float)r.x, ry = (float)r.y, rz = (float)r.z;
    const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    if (r2 < 1e-10f) return;
    const Lambda& lp = J.lp;
    const float a_c6 = (float)la[0], a_c12 = (float)la[1], b_c6 = (float)lb[0], b_c12 = (float)lb[1];
    const float a_q = r0.x * qj, b_q = r0.y * qj, alpha_lj = r0.z, alpha_crf = r0.w;
    const float r4 = r2 * r2, r6 = r4 * r2;
    // hardware reciprocal and reciprocal square root (about 1 ulp; this share
    // is FP32 arithmetic throughout)
    const float a_c126 = fabsf(a_c6) > 1e-10f ? a_c12 * __builtin_amdgcn_rcpf(a_c6) : 0.0f;
    const float b_c126 = fabsf(b_c6) > 1e-10f ? b_c12 * __builtin_amdgcn_rcpf(b_c6) : 0.0f;
    const float a_d2 = r2 + alpha_crf * (float)lp.b_crfs_lambda2;
    const float b_d2 = r2 + alpha_crf * (float)lp.a_crfs_lambda2;
    const float a_di = __builtin_amdgcn_rsqf(a_d2), b_di = __builtin_amdgcn_rsqf(b_d2);
    const float a_d3i = a_di * a_di * a_di, b_d3i = b_di * b_di * b_di;
    const float a_d6i = __builtin_amdgcn_rcpf(r6 + alpha_lj * (float)lp.b_ljs_lambda2 * a_c126);
    const float b_d6i = __builtin_amdgcn_rcpf(r6 + alpha_lj * (float)lp.a_ljs_lambda2 * b_c126);
    const float aln = (float)lp.a_lj_lambda_n, bln = (float)lp.b_lj_lambda_n;
    float f = ((float)lp.a_crf_lambda_n * a_q * (a_d3i + r1.x) + (float)lp.b_crf_lambda_n * b_q * (b_d3i + r1.y)) *
              (float)J.fpepsi;
    f += -6.0f * (aln * a_c6 * a_d6i * a_d6i + bln * b_c6 * b_d6i * b_d6i) * r4;
    f += 12.0f * (aln * a_c12 * a_d6i * a_d6i * a_d6i + bln * b_c12 * b_d6i * b_d6i * b_d6i) * r4;
    fx -= (double)(rx * f);
    fy -= (double)(ry * f);
    fz -= (double)(rz * f);
}

// perturbed_excluded_correction for one excluded pair i < j (perturbed.rs:445-510)
__device__ __forceinline__ void pert_excl_delta(const SysPtrs& p, const JobDev& J, int32_t i, int32_t j, bool first,
                                                double (&v)[ROW], double& fx, double& fy, double& fz) {
    PertPair q = pert_params(p, J, i, j);
    d3 r = nim(J.box, ld3(p.x_cur, i), ld3(p.x_cur, j));
    double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    PertDelta d = pert_excl64(r2, q.a_q, q.b_q, q.acrf, J.lp, J.crf, J.fpepsi);
    own_side(first, r, d.f, fx, fy, fz);
    if (first) {
        v[1] += d.de_crf;
        v[11] += d.dhdl;
    }
}

// perturbed_one_four_correction for one 1-4 pair i < j (perturbed.rs:531-604)
__device__ __forceinline__ void pert_14_delta(const SysPtrs& p, const JobDev& J, int32_t i, int32_t j, bool first,
                                              double (&v)[ROW], double& fx, double& fy, double& fz) {
    PertPair q = pert_params(p, J, i, j);
    d3 r = nim(J.box, ld3(p.x_cur, i), ld3(p.x_cur, j));
    double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    if (r2 < 1e-10) return;
    PertDelta d = pert_14_64(r.x, r.y, r.z, r2, q.la, q.lb, q.a_q, q.b_q, q.alj, q.acrf, J.lp, J.crf, J.fpepsi);
    own_side(first, r, d.f, fx, fy, fz);
    if (first) {
        v[0] += d.de_lj;
        v[1] += d.de_crf;
        v[11] += d.dhdl;
    }
}

// perturbed_atom_pair_correction for one PERTATOMPAIR entry (perturbed.rs:624-707);
// only the LJ energy delta is added (forcefield.rs:923-928)
__device__ __forceinline__ void pert_atom_pair_delta(const SysPtrs& p, const JobDev& J, const int32_t* ap, bool first,
                                                     double (&v)[ROW], double& fx, double& fy, double& fz) {
    int32_t i = ap[0], j = ap[1], at = ap[2], bt = ap[3];
    bool in14 = false;
    for (int32_t kk = p.o14_off[i]; kk < p.o14_off[i + 1]; ++kk) in14 |= p.o14[kk] == j;
    for (int32_t kk = p.o14_off[j]; kk < p.o14_off[j + 1]; ++kk) in14 |= p.o14[kk] == i;
    if (!in14) return;
    d3 r = nim(J.box, ld3(p.x_cur, i), ld3(p.x_cur, j));
    double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    if (r2 < 1e-10) return;
    const double* lj = p.lj + J.lj_off + 4 * ((int64_t)p.iac[i] * J.n_types + p.iac[j]);
    PertDelta d = pert_atom_pair64(r2, lj, p.q[i] * p.q[j] * J.fpepsi, at, bt, J.lp, J.crf);
    own_side(first, r, d.f, fx, fy, fz);
    if (first) {
        v[0] += d.de_lj;
        v[11] += d.dhdl;
    }
}

// One pair (owner atom a, partner atom b) in the reference's orientation.
// cg path: shift from the two groups' first atoms (process_pairs_cg_grouped,
// innerloops.rs:117-135); otherwise per-atom nearest image (innerloops.rs:180-186).
__device__ __forceinline__ void gen_pair(const SysPtrs& p, const JobDev& J, double (&v)[ROW],
                                         double& fx, double& fy, double& fz, bool canon, bool cgpath,
                                         d3 pa, d3 pb, d3 fa0, d3 fb0, int32_t ta, int32_t tb,
                                         double qa, double qb, bool virial) {
    d3 r;
    if (cgpath) {
        d3 A0 = canon ? fa0 : fb0, B0 = canon ? fb0 : fa0;
        d3 rf = nim(J.box, A0, B0);
        double tx = rf.x - A0.x + B0.x, ty = rf.y - A0.y + B0.y, tz = rf.z - A0.z + B0.z;
        d3 u = canon ? pa : pb, w = canon ? pb : pa;
        r = d3{u.x + tx - w.x, u.y + ty - w.y, u.z + tz - w.z};
    } else {
        r = canon ? nim(J.box, pa, pb) : nim(J.box, pb, pa);
    }
    int32_t ti = canon ? ta : tb, tj = canon ? tb : ta;
    const double* lj = p.lj + J.lj_off + 4 * ((int64_t)ti * J.n_types + tj);
    double qi = canon ? qa : qb, qj = canon ? qb : qa;
    double qp = qi * qj * J.fpepsi;
    double fm, elj, ecrf;
    lj_crf64(r.x, r.y, r.z, lj[0], lj[1], qp, J.crf, fm, elj, ecrf);
    double gx = r.x * fm, gy = r.y * fm, gz = r.z * fm;
    if (canon) {
        fx += gx; fy += gy; fz += gz;
        v[0] += elj;
        v[1] += ecrf;
        if (virial) add_virial(v, r, gx, gy, gz);
    } else {
        fx -= gx; fy -= gy; fz -= gz;
    }
}

// The pairs of one sorted atom a (original o) from its group's general list
// and, for solute groups, its own group: entries e0 + lane0, e0 + lane0 + stride,
// ... (one lane: lane0 = 0, stride = 1; a wave: lane0 = lane, stride = 64).
template <bool LONG, bool SOLV>
__device__ __forceinline__ void gen_atom64(const SysPtrs& p, int32_t a, int32_t o, int32_t lane0, int32_t stride,
                                           int32_t virial, double (&v)[ROW], double& fx, double& fy, double& fz) {
    int32_t k = p.s_atom_cg[a];
    const JobDev& J = p.jobs[p.atom_job[o]];
    int32_t sk = p.s_cg_start[k];
    int32_t nk = p.s_cg_n[k];
    double4 xa = p.xq[a];
    d3 pa = xyz(xa);
    d3 pk0 = xyz(p.xq[sk]);
    int32_t ta = p.s_type[a];
    bool cgp = J.use_cg_grouped != 0;
    const bool jp = !LONG && J.has_pert;   // perturbation deltas ride on the short pass
    const bool po = !SOLV && jp && p.pflag[o];   // solvent atoms are never perturbed
    PertAtom PA{};
    if (po) PA = pert_atom(p, J, o);
    // entries written: a group that outgrew its slot has only cap of them until the refill
    int32_t e0 = (int32_t)p.gen_off[k], e1 = e0 + (int32_t)min(p.gen_cnt[k], p.gen_cap[k]);
    for (int32_t e = e0 + lane0; e < e1; e += stride) {
        uint32_t ent = p.gen_ent[e];
        bool lng = (ent & ENT_LONG) != 0;
        int32_t m = (int32_t)(ent & ENT_IDX);
        bool canon = (ent & ENT_CANON) != 0;
        // the pair's first atom lies in the lower group; only its group can
        // hold the perturbed first atom (perturbed.rs:327-332)
        bool pchk = jp && p.cg_pert[p.s_cg_orig[canon ? k : m]];
        int32_t sm = p.s_cg_start[m], nm = p.s_cg_n[m];
        if (lng != LONG) {
            // long pairs are cached from the last rebuild; their correction is
            // taken every step at the current positions (forcefield.rs:796-830)
            if (!pchk) continue;
            for (int32_t b = 0; b < nm; ++b) {
                int32_t ob = p.s2o[sm + b];
                int32_t fi = canon ? o : ob, fj = canon ? ob : o;
                if (p.pflag[fi]) pert_pair_delta(p, J, fi, fj, canon, v, fx, fy, fz);
            }
            continue;
        }
        bool ss = (ent & ENT_SS) != 0;
        d3 pm0 = xyz(p.xq[sm]);
        // a perturbed first atom with an unperturbed partner group: partner data
        // from the sorted copies
        const bool pfast = po && canon && !p.cg_pert[p.s_cg_orig[m]];
        for (int32_t b = 0; b < nm; ++b) {
            if (!LONG && ss) {
                int32_t ob = p.s2o[sm + b];
                if (canon ? excluded_or_14(p, o, ob) : excluded_or_14(p, ob, o)) continue;
            }
            if (pfast) {
                pert_pair_direct_s(p, J, PA, p.xq[sm + b], p.s_type[sm + b], virial != 0, v, fx, fy, fz);
                continue;
            }
            if (pchk) {
                int32_t ob = p.s2o[sm + b];
                if (canon && po) {
                    pert_pair_direct_i(p, J, PA, ob, virial != 0, v, fx, fy, fz);
                    continue;
                }
                if (!canon && p.pflag[ob]) {
                    pert_pair_direct(p, J, ob, o, false, virial != 0, v, fx, fy, fz);
                    continue;
                }
            }
            double4 xb = p.xq[sm + b];
            gen_pair(p, J, v, fx, fy, fz, canon, !LONG && cgp, pa, xyz(xb), pk0, pm0, ta,
                     p.s_type[sm + b], xa.w, xb.w, LONG || virial);
        }
    }
    // intra-group pairs of solute groups (pairlist.rs:238-249), always short
    if (!LONG && p.s_cg_orig[k] - J.cg_off < J.n_solute_cg) {
        int32_t ia = a - sk;
        bool pown = jp && p.cg_pert[p.s_cg_orig[k]];
        for (int32_t b = lane0; b < nk; b += stride) {
            if (b == ia) continue;
            int32_t ob = p.s2o[sk + b];
            bool canon = ia < b;
            if (canon ? excluded_or_14(p, o, ob) : excluded_or_14(p, ob, o)) continue;
            if (pown) {
                if (canon && po) {
                    pert_pair_direct_i(p, J, PA, ob, virial != 0, v, fx, fy, fz);
                    continue;
                }
                if (!canon && p.pflag[ob]) {
                    pert_pair_direct(p, J, ob, o, false, virial != 0, v, fx, fy, fz);
                    continue;
                }
            }
            double4 xb = p.xq[sk + b];
            gen_pair(p, J, v, fx, fy, fz, canon, cgp, pa, xyz(xb), pk0, pk0, ta, p.s_type[sk + b],
                     xa.w, xb.w, virial);
        }
    }
}

// Every non-solvent-solvent-short pair: solute-solute (with exclusions), solute-
// solvent, intra-group, and (LONG) the twin-range long pairs; one lane per
// solvent atom (a solvent atom has few such partners).  Solute atoms, whose
// lists are long, are done a wave each by k_gen_sol.
template <bool LONG>
__global__ void __launch_bounds__(BLOCK) k_gen(SysPtrs p, int32_t npad, int32_t virial) {
    double v[ROW] = {0};
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t o = a < npad ? p.s2o[a] : -1;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (o >= 0 && !held(p, p.atom_job[o])) {
        const JobDev& J = p.jobs[p.atom_job[o]];
        if (o - J.atom_off >= J.n_solute_atoms) gen_atom64<LONG, true>(p, a, o, 0, 1, virial, v, fx, fy, fz);
    }
    if (a < npad) {
        double* f = LONG ? p.f_long_s : p.f_s;
        if (LONG) {
            f[3 * a] = fx;
            f[3 * a + 1] = fy;
            f[3 * a + 2] = fz;
        } else {
            f[3 * a] += fx;
            f[3 * a + 1] += fy;
            f[3 * a + 2] += fz;
        }
    }
    write_row(v, p.rows, LONG ? SLOT_LONG : SLOT_GEN, p.nrow_blocks);
}


// Mixed-precision counterpart of gen_atom64: FP32 pair arithmetic on offsets
// from the groups' first atoms (their nearest-image vector taken in FP64), FP64
// sums; pairs with a perturbed first atom entirely in FP64.  Energies from the
// canonical partner only.
template <bool LONG, bool VIR, bool SOLV>
__device__ __forceinline__ void gen_atom32(const SysPtrs& p, int32_t a, int32_t o, int32_t lane0, int32_t stride,
                                           double (&v)[ROW], double& fx, double& fy, double& fz) {
    int32_t k = p.s_atom_cg[a];
    const JobDev& J = p.jobs[p.atom_job[o]];
    const Box box = J.box;
    const float c2 = (float)J.crf.crf_2cut3i, cc = (float)J.crf.crf_cut, c3 = (float)J.crf.crf_cut3i;
    const float fpe = (float)J.fpepsi;
    const int32_t nt = J.n_types;
    const double* ljb = p.lj + J.lj_off;
    int32_t sk = p.s_cg_start[k], nk = p.s_cg_n[k];
    float4 da = p.xf[a];
    double4 xa = p.xq[a];
    d3 pk0 = xyz(p.xq[sk]);
    int32_t ta = p.s_type[a];
    bool cgp = J.use_cg_grouped != 0 && !LONG;
    const bool jp = !LONG && J.has_pert;
    // SOLV (a solvent atom's lane): solvent atoms are never perturbed, so the
    // lane never evaluates a perturbed pair (its first atom's lanes do both
    // shares) and carries no FP64 soft-core code
    const bool po = !SOLV && jp && p.pflag[o];
    PertAtom PA{};
    if (po) PA = pert_atom(p, J, o);
    // entries written: a group that outgrew its slot has only cap of them until the refill
    int32_t e0 = (int32_t)p.gen_off[k], e1 = e0 + (int32_t)min(p.gen_cnt[k], p.gen_cap[k]);
    // software pipeline over this lane's entries: the entry word two ahead and the
    // partner range one ahead are loaded while the current entry is evaluated
    // (the loop is a chain of dependent loads otherwise)
    int32_t e = e0 + lane0;
    uint32_t ent1 = e < e1 ? p.gen_ent[e] : 0u;
    uint32_t ent2 = e + stride < e1 ? p.gen_ent[e + stride] : 0u;
    int32_t sm1 = 0, nm1 = 0;
    if (e < e1) {
        sm1 = p.s_cg_start[ent1 & ENT_IDX];
        nm1 = p.s_cg_n[ent1 & ENT_IDX];
    }
    for (; e < e1; e += stride) {
        const uint32_t ent = ent1;
        const int32_t sm = sm1, nm = nm1;
        ent1 = ent2;
        if (e + stride < e1) {
            sm1 = p.s_cg_start[ent1 & ENT_IDX];
            nm1 = p.s_cg_n[ent1 & ENT_IDX];
        }
        if (e + 2 * stride < e1) ent2 = p.gen_ent[e + 2 * stride];
        bool lng = (ent & ENT_LONG) != 0;
        int32_t m = (int32_t)(ent & ENT_IDX);
        float canon = (ent & ENT_CANON) ? 1.0f : 0.0f;
        bool pchk = !SOLV && jp && p.cg_pert[p.s_cg_orig[canon != 0.f ? k : m]];
        if (lng != LONG) {
            // cached long pairs: the correction is taken every step at the current
            // positions (forcefield.rs:796-830), each atom its own share (FP64)
            if (SOLV || !pchk) continue;
            for (int32_t b = 0; b < nm; ++b) {
                int32_t ob = p.s2o[sm + b];
                int32_t fi = canon != 0.f ? o : ob, fj = canon != 0.f ? ob : o;
                if (p.pflag[fi]) pert_pair_delta(p, J, fi, fj, canon != 0.f, v, fx, fy, fz);
            }
            continue;
        }
        bool ss = (ent & ENT_SS) != 0;
        float sx = 0.f, sy = 0.f, sz = 0.f;
        if (cgp) shift32(box, pk0, xyz(p.xq[sm]), sx, sy, sz);
        // a perturbed first atom with an unperturbed partner group: partner data
        // from the sorted copies
        const bool pfast = po && canon != 0.f && !p.cg_pert[p.s_cg_orig[m]];
        for (int32_t b = 0; b < nm; ++b) {
            if (!LONG && ss) {
                int32_t ob = p.s2o[sm + b];
                if (canon != 0.f ? excluded_or_14(p, o, ob) : excluded_or_14(p, ob, o)) continue;
            }
            double4 xb = p.xq[sm + b];
            if (SOLV && jp && canon == 0.f) {
                // a perturbed partner (always the pair's first atom): this lane's
                // FP32 share
                const int32_t w = p.s_prec[sm + b];
                if (w >= 0) {
                    pert_share32(p, J, w, xyz(xb), xyz(xa), ta, (float)xa.w, fx, fy, fz);
                    continue;
                }
            }
            if (pfast) {
                pert_pair_direct_s(p, J, PA, xb, p.s_type[sm + b], VIR, v, fx, fy, fz);
                continue;
            }
            if (pchk) {
                // a perturbed first atom: its lanes take the soft-core pair in FP64
                // with the energies; the second atom's lanes take their own share
                // (a solvent atom in FP32, a solute atom in FP64)
                int32_t ob = p.s2o[sm + b];
                if (canon == 0.f) {
                    if (p.pflag[ob]) {
                        pert_pair_direct(p, J, ob, o, false, VIR, v, fx, fy, fz);
                        continue;
                    }
                } else if (po) {
                    pert_pair_direct_i(p, J, PA, ob, VIR, v, fx, fy, fz);
                    continue;
                }
            }
            float rx, ry, rz;
            if (cgp) {
                float4 db = p.xf[sm + b];
                rx = (sx + da.x) - db.x;
                ry = (sy + da.y) - db.y;
                rz = (sz + da.z) - db.z;
            } else {
                d3 r = nim(box, xyz(xa), xyz(xb));
                rx = (float)r.x; ry = (float)r.y; rz = (float)r.z;
            }
            const double* lj = ljb + 4 * ((int64_t)ta * nt + p.s_type[sm + b]);
            float fm, elj, ecrf;
            lj_crf32(rx, ry, rz, (float)lj[0], (float)lj[1], (float)xa.w * (float)xb.w * fpe, c2, cc, c3, fm, elj,
                     ecrf);
            float gx = rx * fm, gy = ry * fm, gz = rz * fm;
            fx += (double)gx;
            fy += (double)gy;
            fz += (double)gz;
            v[0] += (double)(canon * elj);
            v[1] += (double)(canon * ecrf);
            if (VIR || LONG)
                add_virial(v, d3{(double)(canon * rx), (double)(canon * ry), (double)(canon * rz)}, gx, gy, gz);
        }
    }
    if (!SOLV && !LONG && p.s_cg_orig[k] - J.cg_off < J.n_solute_cg) {
        int32_t ia = a - sk;
        bool pown = jp && p.cg_pert[p.s_cg_orig[k]];
        for (int32_t b = lane0; b < nk; b += stride) {
            if (b == ia) continue;
            int32_t ob = p.s2o[sk + b];
            bool canon = ia < b;
            if (canon ? excluded_or_14(p, o, ob) : excluded_or_14(p, ob, o)) continue;
            double4 xb = p.xq[sk + b];
            if (pown) {
                if (canon && po) {
                    pert_pair_direct_i(p, J, PA, ob, VIR, v, fx, fy, fz);
                    continue;
                }
                if (!canon && p.pflag[ob]) {
                    pert_pair_direct(p, J, ob, o, false, VIR, v, fx, fy, fz);
                    continue;
                }
            }
            float rx = (float)(xa.x - xb.x), ry = (float)(xa.y - xb.y), rz = (float)(xa.z - xb.z);
            const double* lj = ljb + 4 * ((int64_t)ta * nt + p.s_type[sk + b]);
            float fm, elj, ecrf;
            lj_crf32(rx, ry, rz, (float)lj[0], (float)lj[1], (float)xa.w * (float)xb.w * fpe, c2, cc, c3, fm, elj,
                     ecrf);
            float gx = rx * fm, gy = ry * fm, gz = rz * fm;
            fx += (double)gx;
            fy += (double)gy;
            fz += (double)gz;
            float cf = canon ? 1.0f : 0.0f;
            v[0] += (double)(cf * elj);
            v[1] += (double)(cf * ecrf);
            if (VIR) add_virial(v, d3{(double)(cf * rx), (double)(cf * ry), (double)(cf * rz)}, gx, gy, gz);
        }
    }
}

template <bool LONG, bool VIR>
__global__ void __launch_bounds__(BLOCK) k_gen_mx(SysPtrs p, int32_t npad) {
    double v[ROW] = {0};
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t o = a < npad ? p.s2o[a] : -1;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (o >= 0 && !held(p, p.atom_job[o])) {
        const JobDev& J = p.jobs[p.atom_job[o]];
        if (o - J.atom_off >= J.n_solute_atoms) gen_atom32<LONG, VIR, true>(p, a, o, 0, 1, v, fx, fy, fz);
    }
    if (a < npad) {
        long long ix = to_fx(fx), iy = to_fx(fy), iz = to_fx(fz);
        if (LONG) {
            p.fi_long_s[3 * a] = ix;
            p.fi_long_s[3 * a + 1] = iy;
            p.fi_long_s[3 * a + 2] = iz;
        } else {
            p.fi_s[3 * a] += ix;
            p.fi_s[3 * a + 1] += iy;
            p.fi_s[3 * a + 2] += iz;
        }
    }
    write_row(v, p.rows, LONG ? SLOT_LONG : SLOT_GEN, p.nrow_blocks);
}

// Solvent atoms' short general pairs in mixed precision, only for the solvent
// groups listed by the pair search (gsol): GSV_SLOT lanes per group (one group
// per slot, GSV_LANES lanes per atom taking every GSV_LANES-th entry), FP64 lane
// sums combined by a fixed butterfly and added to the atom's fixed-point force
// after the solvent pair kernel has written it.  A solvent group's short
// general entries are its solute partners (solvent-solvent short pairs are all
// in the solvent list), and the solute atom comes first in the topology, so
// those pairs' energies and virial belong to the solute atom's block
// (k_gen_sol): these lanes add none.  (Run beside the solvent pair kernel into
// a per-group buffer added afterwards, it measured slower: 2.25e11 against
// 2.33e11 pairs/s on the 64-job FEP mix.)
template <int GSV_LANES>
__global__ void __launch_bounds__(BLOCK) k_gen_sv(SysPtrs p, int32_t ngsol) {
    constexpr int GSV_SLOT = 4 * GSV_LANES;   // solvent molecules have at most 4 atoms
    const int32_t w = ((int32_t)blockIdx.x * BLOCK + (int32_t)threadIdx.x) / GSV_SLOT;
    const int32_t r = (int32_t)threadIdx.x % GSV_SLOT, i = r / GSV_LANES, c = r % GSV_LANES;
    double v[ROW] = {0};
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int32_t a = -1;
    if (w < ngsol) {
        const int32_t k = p.gsol[w];
        const int32_t sk = p.s_cg_start[k];
        if (i < p.s_cg_n[k] && !held(p, p.atom_job[p.s2o[sk]])) a = sk + i;
    }
    if (a >= 0) gen_atom32<false, false, true>(p, a, p.s2o[a], c, GSV_LANES, v, fx, fy, fz);
#pragma unroll
    for (int m = 1; m < GSV_LANES; m <<= 1) {
        fx += __shfl_xor(fx, m, WAVE);
        fy += __shfl_xor(fy, m, WAVE);
        fz += __shfl_xor(fz, m, WAVE);
    }
    if (c == 0 && a >= 0) {
        p.fi_s[3 * a] += to_fx(fx);
        p.fi_s[3 * a + 1] += to_fx(fy);
        p.fi_s[3 * a + 2] += to_fx(fz);
    }
}

// Solute atoms: one block of NT lanes per atom (global solute index w, original
// atom sol_atom[w]), lane t taking every NT-th entry from t; forces and energies
// summed over each wave with a fixed xor butterfly and over the waves in order.
// Energies go to the per-solute-atom record sol_e[(LONG ? nsol : 0) + w] that
// k_reduce adds per job.  (Eight waves per atom measured slower than four on
// the 64-job FEP mix: 60 against 46 us per step.)
template <bool LONG, bool MIXED, bool VIR, int NT>
__global__ void __launch_bounds__(NT) k_gen_sol(SysPtrs p, int32_t nsol, int32_t virial) {
    constexpr int NV = ROW + 3;
    __shared__ double lds[NT / WAVE][NV];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const int32_t w = blockIdx.x;
    double v[ROW] = {0};
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t o = p.sol_atom[w];
    if (held(p, p.atom_job[o])) return;   // block-uniform, before any barrier
    const int32_t a = p.o2s[o];
    if (MIXED) gen_atom32<LONG, VIR, false>(p, a, o, threadIdx.x, NT, v, fx, fy, fz);
    else gen_atom64<LONG, false>(p, a, o, threadIdx.x, NT, virial, v, fx, fy, fz);
#pragma unroll
    for (int m = 1; m < WAVE; m <<= 1) {
        fx += __shfl_xor(fx, m, WAVE);
        fy += __shfl_xor(fy, m, WAVE);
        fz += __shfl_xor(fz, m, WAVE);
#pragma unroll
        for (int c = 0; c < ROW; ++c) v[c] += __shfl_xor(v[c], m, WAVE);
    }
    if (lane == 0) {
        lds[wv][0] = fx;
        lds[wv][1] = fy;
        lds[wv][2] = fz;
#pragma unroll
        for (int c = 0; c < ROW; ++c) lds[wv][3 + c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) t[c] = lds[0][c];
        for (int q = 1; q < NT / WAVE; ++q)
#pragma unroll
            for (int c = 0; c < NV; ++c) t[c] += lds[q][c];
        if (MIXED) {
            long long ix = to_fx(t[0]), iy = to_fx(t[1]), iz = to_fx(t[2]);
            if (LONG) {
                p.fi_long_s[3 * a] = ix;
                p.fi_long_s[3 * a + 1] = iy;
                p.fi_long_s[3 * a + 2] = iz;
            } else {
                // the atom's whole short-range pair force (the solvent pair kernel
                // leaves solute atoms alone): assigned, one block per atom
                p.fi_s[3 * a] = ix;
                p.fi_s[3 * a + 1] = iy;
                p.fi_s[3 * a + 2] = iz;
            }
        } else {
            if (LONG) {
                p.f_long_s[3 * a] = t[0];
                p.f_long_s[3 * a + 1] = t[1];
                p.f_long_s[3 * a + 2] = t[2];
            } else {
                p.f_s[3 * a] += t[0];
                p.f_s[3 * a + 1] += t[1];
                p.f_s[3 * a + 2] += t[2];
            }
        }
        double* r = p.sol_e + ((int64_t)(LONG ? nsol : 0) + w) * ROW;
#pragma unroll
        for (int c = 0; c < ROW; ++c) r[c] = t[3 + c];
    }
}

// RF self / excluded pairs (excluded.rs:25-112) and 1-4 pairs (excluded.rs:131-206),
// one lane per original atom.  Partners below the owner contribute only their
// force, evaluated in the reference's orientation.
// solv_energy = 0 (md steps): the solvent molecules' excluded RF energy is taken
// by the integrator's group lanes (k_integrate), and solvent atoms have nothing
// else here (no force, no 1-4, never perturbed).
__global__ void __launch_bounds__(BLOCK) k_excl14(SysPtrs p, int32_t npad, int32_t solv_energy,
                                                  const int32_t* blocks) {
    double v[ROW] = {0};
    const int32_t blk = blocks ? blocks[blockIdx.x] : (int32_t)blockIdx.x;
    int32_t i = blk * BLOCK + (int32_t)threadIdx.x;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int32_t jb = i < npad ? p.atom_job[i] : -1;
    if (jb >= 0 && !solv_energy && i - p.jobs[jb].atom_off >= p.jobs[jb].n_solute_atoms) jb = -1;
    if (jb >= 0 && held(p, jb)) jb = -1;
    if (jb >= 0) {
        const JobDev& J = p.jobs[jb];
        const double* x = p.x_cur;
        d3 pi = ld3(x, i);
        double qi = p.q[i];
        bool sol_i = i - J.atom_off < J.n_solute_atoms;
        const bool jp = J.has_pert != 0;
        bool pi_ = jp && p.pflag[i];
        if (sol_i) v[1] += -0.5 * qi * qi * J.fpepsi * J.crf.crf_cut;
        if (pi_) {
            // perturbed_self_energy_correction (perturbed.rs:395-425)
            double de, dh;
            pert_self64(p.pa_q[i], p.pb_q[i], J.lp, J.crf, J.fpepsi, de, dh);
            v[1] += de;
            v[11] += dh;
        }
        for (int32_t k = p.excl_off[i]; k < p.excl_off[i + 1]; ++k) {
            int32_t j = p.excl[k];
            d3 pj = ld3(x, j);
            double qj = p.q[j];
            if (j > i) {
                d3 r = nim(J.box, pi, pj);
                double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
                if (sol_i) {
                    double qp = qi * qj * J.fpepsi;
                    v[1] += qp * (-J.crf.crf_2cut3i * r2 - J.crf.crf_cut);
                    double fs = qp * J.crf.crf_cut3i;
                    double gx = r.x * fs, gy = r.y * fs, gz = r.z * fs;
                    fx += gx; fy += gy; fz += gz;
                    add_virial(v, r, gx, gy, gz);
                    if (jp && (pi_ || p.pflag[j])) pert_excl_delta(p, J, i, j, true, v, fx, fy, fz);
                } else {
                    v[1] += -qi * qj * J.fpepsi * J.crf.crf_2cut3i * r2;
                }
            } else if (j < i && j - J.atom_off < J.n_solute_atoms) {
                d3 r = nim(J.box, pj, pi);
                double fs = qj * qi * J.fpepsi * J.crf.crf_cut3i;
                fx -= r.x * fs; fy -= r.y * fs; fz -= r.z * fs;
                if (jp && (pi_ || p.pflag[j])) pert_excl_delta(p, J, j, i, false, v, fx, fy, fz);
            }
        }
        for (int32_t k = p.o14s_off[i]; k < p.o14s_off[i + 1]; ++k) {
            uint32_t ent = (uint32_t)p.o14s[k];
            int32_t j = (int32_t)(ent & 0x7FFFFFFFu);
            bool canon = (ent & 0x80000000u) != 0;
            d3 pj = ld3(x, j);
            d3 r = canon ? nim(J.box, pi, pj) : nim(J.box, pj, pi);
            double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            if (r2 < 1e-10) continue;
            int32_t lo = canon ? i : j, hi = canon ? j : i;
            const double* lj = p.lj + J.lj_off + 4 * ((int64_t)p.iac[lo] * J.n_types + p.iac[hi]);
            double inv_r2 = 1.0 / r2;
            double inv_r6 = inv_r2 * inv_r2 * inv_r2;
            double elj = (lj[3] * inv_r6 - lj[2]) * inv_r6;
            double flj = (12.0 * lj[3] * inv_r6 - 6.0 * lj[2]) * inv_r6 * inv_r2;
            double qp = p.q[lo] * p.q[hi] * J.fpepsi;
            double inv_r = sqrt(inv_r2);
            double ecrf = qp * (inv_r * 1.0 - J.crf.crf_2cut3i * r2 - J.crf.crf_cut);
            double fcrf = qp * (inv_r * 1.0 * inv_r2 + J.crf.crf_cut3i);
            double ft = flj + fcrf;
            double gx = r.x * ft, gy = r.y * ft, gz = r.z * ft;
            if (canon) {
                fx += gx; fy += gy; fz += gz;
                v[0] += elj;
                v[1] += ecrf;
                add_virial(v, r, gx, gy, gz);
            } else {
                fx -= gx; fy -= gy; fz -= gz;
            }
            if (jp && (pi_ || p.pflag[j])) pert_14_delta(p, J, lo, hi, canon, v, fx, fy, fz);
        }
        if (jp) {
            // PERTATOMPAIR entries of this atom (bit 31: this atom is the entry's i)
            for (int32_t k = p.pap_off[i]; k < p.pap_off[i + 1]; ++k) {
                uint32_t ent = (uint32_t)p.pap[k];
                const int32_t* ap = p.papairs + 4 * (int64_t)(ent & 0x7FFFFFFFu);
                pert_atom_pair_delta(p, J, ap, (ent & 0x80000000u) != 0, v, fx, fy, fz);
            }
        }
    }
    if (i < npad && (solv_energy || jb >= 0)) {
        p.f_x[3 * i] = fx;
        p.f_x[3 * i + 1] = fy;
        p.f_x[3 * i + 2] = fz;
    }
    write_row(v, p.rows, SLOT_EXCL, p.nrow_blocks, blk);
}

// total force on each original atom
__global__ void k_assemble(SysPtrs p, int32_t npad, int32_t with_long, int32_t mixed) {
    int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= npad) return;
    int32_t o = p.s2o[a];
    if (o < 0) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double f;
        if (mixed) {
            long long t = p.fi_s[3 * a + c] + p.fi_near[3 * a + c];
            if (with_long) t += p.fi_long_s[3 * a + c];
            f = (double)t * FX_INV;
        } else {
            f = p.f_s[3 * a + c];
            if (with_long) f += p.f_long_s[3 * a + c];
        }
        f += p.f_x[3 * o + c];
        p.f[3 * o + c] = f;
    }
}

// ---------------------------------------------------------------------------
// step
// ---------------------------------------------------------------------------

__device__ __forceinline__ double berendsen_scale(const JobDev& J, double ekin) {
    if (J.tau_t < 0.0) return 1.0;   // no coupling
    const double kB = 0.00831441;    // units.rs:47
    double free_temp = J.dof_t > 0.0 ? 2.0 * ekin / (J.dof_t * kB) : 0.0;
    if (free_temp < 1e-10) free_temp = J.t_ref;
    if (free_temp < 1e-10) return 1.0;
    if (J.tau_t == 0.0) return sqrt(J.t_ref / free_temp);
    return sqrt(1.0 + J.dt / J.tau_t * (J.t_ref / free_temp - 1.0));
}

// NoseHooverThermostat::apply (nosehoover_thermostat.rs:116-227) for one job:
// the chain advanced from its last committed state `z` with E_kin_new into
// `pend` (k_integrate commits it when it applies the scale), and the velocity
// scale, 1 where the reference leaves the velocities untouched
__device__ __forceinline__ double nosehoover_scale(const JobDev& J, const double* z, double* pend, double ekin) {
    const int n = J.nhc;
    for (int i = 0; i < n; ++i) pend[i] = z[i];
    if (J.tau_t < 0.0) return 1.0;   // no coupling
    const double kB = 0.00831441;    // units.rs:47
    const double dt = J.dt, tau = J.tau_t, dof = J.dof_t, t0 = J.t_ref;
    double free_temp = dof > 0.0 ? 2.0 * ekin / (dof * kB) : 0.0;
    if (free_temp < 2.220446049250313e-16) free_temp = t0;   // f64::EPSILON
    if (n == 1) {
        pend[0] = z[0] + dt / (tau * tau) * (free_temp / t0 - 1.0);
    } else {
        // tail to head; tau_vec[0] = tau^2, tau_vec[i > 0] = tau^2 / dof
        const double tau2 = tau * tau, tv1 = tau2 / dof;
        {
            const double zz = z[n - 2];
            pend[n - 1] = z[n - 1] + ((n - 2 == 0 ? tau2 : tv1) * zz * zz - 1.0 / dof) / tv1 * dt;
        }
        for (int i = n - 2; i >= 1; --i) {
            const double zp = z[i - 1];
            pend[i] = z[i] + (((i - 1 == 0 ? tau2 : tv1) * zp * zp - 1.0 / dof) / tv1 - z[i] * pend[i + 1]) * dt;
        }
        pend[0] = z[0] + ((free_temp / t0 - 1.0) / tau2 - pend[1] * z[0]) * dt;
    }
    const double scale = 1.0 - pend[0] * dt;
    return fabs(scale - 1.0) < 1e-15 ? 1.0 : scale;
}

__device__ __forceinline__ double bath_scale(const SysPtrs& p, const JobDev& J, int32_t jb, double ekin) {
    if (J.thermo == 2)
        return nosehoover_scale(J, p.nh_zeta + (int64_t)NHC_MAX * jb,
                                p.nh_pend + (int64_t)NHC_MAX * jb, ekin);
    return berendsen_scale(J, ekin);
}

struct v3 {
    double x, y, z;
};
__device__ __forceinline__ v3 mk(double x, double y, double z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 vdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 vcross(v3 a, v3 b) {  // glam DVec3::cross
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ v3 ldv(const double* p, int64_t i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void stv(double* p, int64_t i, v3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// settle (constraints.rs:815-1036) of one water: old positions po, new
// positions pn and velocities vn updated in place; false when the analytical
// solve fails (sin > 1), pn / vn then unchanged.  FP64 throughout.
__device__ __forceinline__ bool settle_one(const JobDev& J, const v3 (&po)[3], v3 (&pn)[3], v3 (&vn)[3]) {
    double m_o = J.m_o, m_h = J.m_h;
    double half_mo_div_mh = 0.5 * m_o / m_h;
    double rc = 0.5 * J.dist_hh;
    double ra = sqrt(J.dist_oh * J.dist_oh - rc * rc) / (1.0 + half_mo_div_mh);
    double rb = half_mo_div_mh * ra;
    double dt_i = 1.0 / J.dt;
    v3 po0 = po[0], po1 = po[1], po2 = po[2];
    v3 pn0 = pn[0], pn1 = pn[1], pn2 = pn[2];
    v3 b0 = vsub(po1, po0), c0 = vsub(po2, po0);
    v3 d0 = vdiv(vadd(vadd(vmul(pn0, m_o), vmul(pn1, m_h)), vmul(pn2, m_h)), m_o + m_h + m_h);
    v3 a1 = vsub(pn0, d0), b1 = vsub(pn1, d0), c1 = vsub(pn2, d0);
    v3 n0 = vcross(b0, c0);
    v3 n1 = vcross(a1, n0);
    v3 n2 = vcross(n0, n1);
    n0 = vdiv(n0, sqrt(vdot(n0, n0)));
    n1 = vdiv(n1, sqrt(vdot(n1, n1)));
    n2 = vdiv(n2, sqrt(vdot(n2, n2)));
    v3 m1 = mk(n1.x, n2.x, n0.x), m2 = mk(n1.y, n2.y, n0.y), m0 = mk(n1.z, n2.z, n0.z);
    b0 = mk(vdot(n1, b0), vdot(n2, b0), vdot(n0, b0));
    c0 = mk(vdot(n1, c0), vdot(n2, c0), vdot(n0, c0));
    a1 = mk(vdot(n1, a1), vdot(n2, a1), vdot(n0, a1));
    b1 = mk(vdot(n1, b1), vdot(n2, b1), vdot(n0, b1));
    c1 = mk(vdot(n1, c1), vdot(n2, c1), vdot(n0, c1));
    double sinphi = a1.z / ra;
    double omsp2 = 1.0 - sinphi * sinphi;
    if (omsp2 < 0.0) return false;
    double cosphi = sqrt(omsp2);
    double sinpsi = (b1.z - c1.z) / (2.0 * rc * cosphi);
    double omss2 = 1.0 - sinpsi * sinpsi;
    if (omss2 < 0.0) return false;
    double cospsi = sqrt(omss2);
    double minus_rb_cosphi = -rb * cosphi;
    double rc_cospsi = rc * cospsi;
    double rc_sinpsi_sinphi = rc * sinpsi * sinphi;
    double rc_sinpsi_cosphi = rc * sinpsi * cosphi;
    v3 a2 = mk(0.0, ra * cosphi, ra * sinphi);
    v3 b2 = mk(-rc_cospsi, minus_rb_cosphi - rc_sinpsi_sinphi, -rb * sinphi + rc_sinpsi_cosphi);
    v3 c2 = mk(rc_cospsi, minus_rb_cosphi + rc_sinpsi_sinphi, -rb * sinphi - rc_sinpsi_cosphi);
    double alpha = b2.x * (b0.x - c0.x) + b0.y * b2.y + c0.y * c2.y;
    double beta = b2.x * (c0.y - b0.y) + b0.x * b2.y + c0.x * c2.y;
    double gamma = b0.x * b1.y - b1.x * b0.y + c0.x * c1.y - c1.x * c0.y;
    double a2b2 = alpha * alpha + beta * beta;
    double sintheta = (alpha * gamma - beta * sqrt(a2b2 - gamma * gamma)) / a2b2;
    double omst2 = 1.0 - sintheta * sintheta;
    if (omst2 < 0.0) return false;
    double costheta = sqrt(omst2);
    v3 a3 = mk(-a2.y * sintheta, a2.y * costheta, a1.z);
    v3 b3 = mk(b2.x * costheta - b2.y * sintheta, b2.x * sintheta + b2.y * costheta, b1.z);
    v3 c3 = mk(-b2.x * costheta - c2.y * sintheta, -b2.x * sintheta + c2.y * costheta, c1.z);
    v3 pa = vadd(mk(vdot(a3, m1), vdot(a3, m2), vdot(a3, m0)), d0);
    v3 pb = vadd(mk(vdot(b3, m1), vdot(b3, m2), vdot(b3, m0)), d0);
    v3 pc = vadd(mk(vdot(c3, m1), vdot(c3, m2), vdot(c3, m0)), d0);
    v3 da = vsub(pa, pn0), db = vsub(pb, pn1), dc = vsub(pc, pn2);
    pn[0] = pa;
    pn[1] = pb;
    pn[2] = pc;
    vn[0] = vadd(vn[0], vmul(da, dt_i));
    vn[1] = vadd(vn[1], vmul(db, dt_i));
    vn[2] = vadd(vn[2], vmul(dc, dt_i));
    return true;
}

// shake (constraints.rs:213-389) of one solvent molecule held in registers:
// sweeps over the template constraints until a sweep changes nothing (an atom's
// constraints skipped when no constraint moved it in the previous sweep); an
// unconverged constraint (|l^2 - r^2| >= 2 tol l^2) moves both atoms along
// the old bond vector by lambda = (l^2 - r^2) / (2 (old . r) (1/m_i + 1/m_j))
// (shake_one_constraint_full, constraints.rs:501-570); then v = (x_new -
// x_old) / dt for the constrained atoms.  cv += the constraint virial's
// diagonal (old bond^2 lambda / dt^2).  False after 1000 sweeps unconverged.
__device__ __forceinline__ v3 pick4(const v3 (&a)[4], int i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : (i == 2 ? a[2] : a[3]));
}
__device__ __forceinline__ void put4(v3 (&a)[4], int i, v3 v) {
    if (i == 0) a[0] = v;
    if (i == 1) a[1] = v;
    if (i == 2) a[2] = v;
    if (i == 3) a[3] = v;
}
__device__ bool shake_molecule(const JobDev& J, int n, const v3 (&xo)[4], v3 (&xn)[4], v3 (&vn)[4],
                               const double (&im)[4], double (&cv)[3]) {
    const double dt2 = J.dt * J.dt;
    unsigned skip_now = 0u, constrained = 0u;
    bool conv = false;
    int it = 0;
    for (int k = 0; k < J.n_sc; ++k) constrained |= (1u << J.sc_ij[2 * k]) | (1u << J.sc_ij[2 * k + 1]);
    while (it < 1000 && !conv) {
        conv = true;
        ++it;
        unsigned moved = 0u;
        for (int k = 0; k < J.n_sc; ++k) {
            const int i = J.sc_ij[2 * k], j = J.sc_ij[2 * k + 1];
            if ((skip_now >> i & 1u) && (skip_now >> j & 1u)) continue;
            const double imi = i == 0 ? im[0] : (i == 1 ? im[1] : (i == 2 ? im[2] : im[3]));
            const double imj = j == 0 ? im[0] : (j == 1 ? im[1] : (j == 2 ? im[2] : im[3]));
            if (imi == 0.0 && imj == 0.0) continue;
            const double l2 = J.sc_len[k] * J.sc_len[k];
            const v3 ni = pick4(xn, i), nj = pick4(xn, j), oi = pick4(xo, i), oj = pick4(xo, j);
            const v3 r = mk(ni.x - nj.x, ni.y - nj.y, ni.z - nj.z);
            const double diff = l2 - (r.x * r.x + r.y * r.y + r.z * r.z);
            if (fabs(diff) < l2 * J.sh_tol * 2.0) continue;
            const v3 rr = mk(oi.x - oj.x, oi.y - oj.y, oi.z - oj.z);
            const double sp = rr.x * r.x + rr.y * r.y + rr.z * r.z;
            const double ims = imi + imj;
            if (ims < 1e-20) continue;
            const double lambda = diff / (sp * 2.0 * ims);
            const double lo = lambda / dt2;
            cv[0] += rr.x * rr.x * lo;
            cv[1] += rr.y * rr.y * lo;
            cv[2] += rr.z * rr.z * lo;
            const v3 cr = mk(rr.x * lambda, rr.y * lambda, rr.z * lambda);
            put4(xn, i, mk(ni.x + cr.x * imi, ni.y + cr.y * imi, ni.z + cr.z * imi));
            put4(xn, j, mk(nj.x - cr.x * imj, nj.y - cr.y * imj, nj.z - cr.z * imj));
            moved |= (1u << i) | (1u << j);
            conv = false;
        }
        skip_now = ~moved;
    }
    const double dti = 1.0 / J.dt;
    for (int a = 0; a < n; ++a)
        if (constrained >> a & 1u) {
            const v3 q = pick4(xn, a), o = pick4(xo, a);
            put4(vn, a, mk((q.x - o.x) * dti, (q.y - o.y) * dti, (q.z - o.z) * dti));
        }
    return conv;
}

// lincs (constraints.rs:1241-1390) of a job's solute constraints, one thread:
// B[k] = (x_old[j] - x_old[i]) / |...| (zero below 1e-10 nm); rhs = sol =
// sdiag (B . (x[j] - x[i]) - l), then order times rhs' = A rhs (A[k][c] =
// coef (B[k] . B[c]) over the coupled c, ascending), sol += rhs'; x[i] += B
// sdiag sol / m_i, x[j] -= B sdiag sol / m_j; the same again with rhs = sdiag
// (l - sqrt(max(2 l^2 - |x[j] - x[i]|^2, 0))) (the rotational correction);
// then v = (x - x_old) / dt for the constrained atoms.  Work rows per
// constraint in p.lc_work: B (3), the two right-hand sides, the solution.
__device__ void lincs_solute(const SysPtrs& p, const JobDev& J) {
    const int32_t nc = J.sh_n;
    const int32_t* ij = p.sh_ij + 2 * (int64_t)J.sh_off;
    const double* len = p.sh_len + J.sh_off;
    const double* sd = p.lc_sdiag + J.sh_off;
    const int32_t* cp = p.lc_cpl + 2 * (int64_t)J.sh_off;
    double* w = p.lc_work + 6 * (int64_t)J.sh_off;
    double* xn = p.x_new;
    const double* xo = p.x_cur;
    for (int32_t k = 0; k < nc; ++k) {
        const d3 oi = ld3(xo, ij[2 * k]), oj = ld3(xo, ij[2 * k + 1]);
        const double rx = oj.x - oi.x, ry = oj.y - oi.y, rz = oj.z - oi.z;
        const double rl = sqrt(rx * rx + ry * ry + rz * rz);
        w[6 * k] = rl < 1e-10 ? 0.0 : rx / rl;
        w[6 * k + 1] = rl < 1e-10 ? 0.0 : ry / rl;
        w[6 * k + 2] = rl < 1e-10 ? 0.0 : rz / rl;
    }
    for (int pass = 0; pass < 2; ++pass) {
        for (int32_t k = 0; k < nc; ++k) {
            const d3 ni = ld3(xn, ij[2 * k]), nj = ld3(xn, ij[2 * k + 1]);
            const double rx = nj.x - ni.x, ry = nj.y - ni.y, rz = nj.z - ni.z;
            double rhs;
            if (pass == 0) {
                const double proj = w[6 * k] * rx + w[6 * k + 1] * ry + w[6 * k + 2] * rz;
                rhs = sd[k] * (proj - len[k]);
            } else {
                const double diff = 2.0 * len[k] * len[k] - (rx * rx + ry * ry + rz * rz);
                const double pp = diff > 0.0 ? sqrt(diff) : 0.0;
                rhs = sd[k] * (len[k] - pp);
            }
            w[6 * k + 3] = rhs;
            w[6 * k + 5] = rhs;
        }
        int ws = 1;
        for (int rec = 0; rec < J.lincs_u; ++rec) {
            for (int32_t k = 0; k < nc; ++k) {
                double acc = 0.0;
                for (int32_t e = cp[2 * k]; e < cp[2 * k] + cp[2 * k + 1]; ++e) {
                    const int32_t c = p.lc_cidx[e];
                    const double bd = w[6 * k] * w[6 * c] + w[6 * k + 1] * w[6 * c + 1] + w[6 * k + 2] * w[6 * c + 2];
                    acc += p.lc_coef[e] * bd * w[6 * c + 3 + (1 - ws)];
                }
                w[6 * k + 3 + ws] = acc;
                w[6 * k + 5] += acc;
            }
            ws = 1 - ws;
        }
        for (int32_t k = 0; k < nc; ++k) {
            const int64_t i = ij[2 * k], j = ij[2 * k + 1];
            const double corr = sd[k] * w[6 * k + 5];
            const double si = corr * p.inv_mass[i], sj = corr * p.inv_mass[j];
            for (int c = 0; c < 3; ++c) {
                xn[3 * i + c] = xn[3 * i + c] + w[6 * k + c] * si;
                xn[3 * j + c] = xn[3 * j + c] - w[6 * k + c] * sj;
            }
        }
    }
    const double dti = 1.0 / J.dt;
    double* vn = p.v_new;
    for (int32_t k = 0; k < nc; ++k)
        for (int s = 0; s < 2; ++s) {
            const int64_t a = ij[2 * k + s];
            for (int c = 0; c < 3; ++c) vn[3 * a + c] = (xn[3 * a + c] - xo[3 * a + c]) * dti;
        }
}

// a job's solute constraints after k_integrate, a block per job: by LINCS
// (lincs_solute) when the job asks for it, else by shake (constraints.rs:213-389,
// the solute's list in the caller's order): one thread sweeps (the constraints
// couple the solute's charge groups), skip flags in p.sh_skip; then v = (x_new -
// x_old) / dt for the constrained atoms and, for a pressure-coupled job, the
// constraint virial into the first group's row.  Either way the solute groups'
// E_kin terms again (as k_integrate writes them).
__global__ void __launch_bounds__(BLOCK) k_shake_solute(SysPtrs p) {
    const int32_t jb = blockIdx.x;
    const JobDev& J = p.jobs[jb];
    if (J.sh_n == 0) return;
    const int32_t step = *p.step;
    if (step >= J.stop) return;
    const int32_t* ij = p.sh_ij + 2 * (int64_t)J.sh_off;
    const double* len = p.sh_len + J.sh_off;
    double* xn = p.x_new;
    const double* xo = p.x_cur;
    if (threadIdx.x == 0 && J.lincs_u > 0) {
        lincs_solute(p, J);
    } else if (threadIdx.x == 0) {
        uint8_t* now = p.sh_skip + 2 * (int64_t)J.atom_off;   // two flags per atom
        for (int32_t a = 0; a < J.n_solute_atoms; ++a) {
            now[2 * a] = 0;
            now[2 * a + 1] = 1;
        }
        const double dt2 = J.dt * J.dt;
        double cv[3] = {0.0, 0.0, 0.0};
        bool conv = false;
        int it = 0, cur = 0;
        while (it < 1000 && !conv) {
            conv = true;
            ++it;
            for (int32_t k = 0; k < J.sh_n; ++k) {
                const int32_t i = ij[2 * k], j = ij[2 * k + 1];
                const int32_t li = i - J.atom_off, lj = j - J.atom_off;
                if (now[2 * li + cur] && now[2 * lj + cur]) continue;
                const double imi = p.inv_mass[i], imj = p.inv_mass[j];
                if (imi == 0.0 && imj == 0.0) continue;
                const double l2 = len[k] * len[k];
                const d3 ni = ld3(xn, i), nj = ld3(xn, j), oi = ld3(xo, i), oj = ld3(xo, j);
                const double rx = ni.x - nj.x, ry = ni.y - nj.y, rz = ni.z - nj.z;
                const double diff = l2 - (rx * rx + ry * ry + rz * rz);
                if (fabs(diff) < l2 * J.sh_tol * 2.0) continue;
                const double ax = oi.x - oj.x, ay = oi.y - oj.y, az = oi.z - oj.z;
                const double sp = ax * rx + ay * ry + az * rz;
                const double ims = imi + imj;
                if (ims < 1e-20) continue;
                const double lambda = diff / (sp * 2.0 * ims);
                const double lo = lambda / dt2;
                cv[0] += ax * ax * lo;
                cv[1] += ay * ay * lo;
                cv[2] += az * az * lo;
                const double cx = ax * lambda, cy = ay * lambda, cz = az * lambda;
                xn[3 * (int64_t)i] = ni.x + cx * imi;
                xn[3 * (int64_t)i + 1] = ni.y + cy * imi;
                xn[3 * (int64_t)i + 2] = ni.z + cz * imi;
                xn[3 * (int64_t)j] = nj.x - cx * imj;
                xn[3 * (int64_t)j + 1] = nj.y - cy * imj;
                xn[3 * (int64_t)j + 2] = nj.z - cz * imj;
                now[2 * li + (1 - cur)] = 0;
                now[2 * lj + (1 - cur)] = 0;
                conv = false;
            }
            // the next sweep's flags become the current ones; the others start set
            cur = 1 - cur;
            for (int32_t a = 0; a < J.n_solute_atoms; ++a) now[2 * a + (1 - cur)] = 1;
        }
        if (!conv) atomicMin(p.settle_fail + jb, step);
        const double dti = 1.0 / J.dt;
        double* vn = p.v_new;
        for (int32_t k = 0; k < J.sh_n; ++k)
            for (int s = 0; s < 2; ++s) {
                const int64_t a = ij[2 * k + s];
                for (int c = 0; c < 3; ++c) vn[3 * a + c] = (xn[3 * a + c] - xo[3 * a + c]) * dti;
            }
        if (J.baro)
            for (int c = 0; c < 3; ++c) p.cg_cvir[3 * (int64_t)J.cg_off + c] += cv[c];
    }
    __syncthreads();
    // the solute groups' E_kin terms from the constrained velocities (k_integrate's order)
    for (int32_t c = J.cg_off + threadIdx.x; c < J.cg_off + J.n_solute_cg; c += blockDim.x) {
        double ek = 0.0, ekn = 0.0;
        for (int32_t i = p.cg_start[c]; i < p.cg_start[c] + p.cg_n[c]; ++i) {
            const d3 vn = ld3(p.v_new, i), vo = ld3(p.v_cur, i);
            const double a2 = vn.x * vn.x + vn.y * vn.y + vn.z * vn.z;
            const double b2 = vo.x * vo.x + vo.y * vo.y + vo.z * vo.z;
            ek += 0.5 * p.mass[i] * (a2 + b2) / 2.0;
            ekn += 0.5 * p.mass[i] * a2;
        }
        p.cg_ek[3 * (int64_t)c] = ek;
        p.cg_ek[3 * (int64_t)c + 1] = ekn;
    }
}

// one LINCS expansion of a solvent molecule's template (at most 6 constraints,
// unrolled so the rows stay in registers): dst[k] = sum over the coupled c
// (ascending) of coef[k][c] (B[k] . B[c]) src[c]; sol += dst
__device__ __forceinline__ void lincs_expand(const JobDev& J, int nc, const v3 (&b)[6], const double (&src)[6],
                                             double (&dst)[6], double (&sol)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k >= nc) continue;
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            if (c >= nc || !(J.lc_cpl[k] >> c & 1)) continue;
            const double bd = b[k].x * b[c].x + b[k].y * b[c].y + b[k].z * b[c].z;
            acc += J.lc_coef[6 * k + c] * bd * src[c];
        }
        dst[k] = acc;
        sol[k] += acc;
    }
}

// lincs of one solvent molecule by the template (lincs_buffered,
// constraints.rs:1426-1440; the same steps as lincs_solute) in registers
__device__ void lincs_molecule(const JobDev& J, const v3 (&xo)[4], v3 (&xn)[4], const double (&im)[4],
                               unsigned& constrained) {
    const int nc = J.n_sc;
    v3 b[6];
    double r0[6], r1[6], sol[6];
    constrained = 0u;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        b[k] = mk(0.0, 0.0, 0.0);
        r0[k] = r1[k] = sol[k] = 0.0;
        if (k >= nc) continue;
        const int i = J.sc_ij[2 * k], j = J.sc_ij[2 * k + 1];
        constrained |= (1u << i) | (1u << j);
        const v3 oi = pick4(xo, i), oj = pick4(xo, j);
        const double rx = oj.x - oi.x, ry = oj.y - oi.y, rz = oj.z - oi.z;
        const double rl = sqrt(rx * rx + ry * ry + rz * rz);
        if (!(rl < 1e-10)) b[k] = mk(rx / rl, ry / rl, rz / rl);
    }
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k >= nc) continue;
            const v3 ni = pick4(xn, J.sc_ij[2 * k]), nj = pick4(xn, J.sc_ij[2 * k + 1]);
            const double rx = nj.x - ni.x, ry = nj.y - ni.y, rz = nj.z - ni.z;
            const double l = J.sc_len[k];
            if (pass == 0) {
                r0[k] = J.lc_sdiag[k] * ((b[k].x * rx + b[k].y * ry + b[k].z * rz) - l);
            } else {
                const double diff = 2.0 * l * l - (rx * rx + ry * ry + rz * rz);
                r0[k] = J.lc_sdiag[k] * (l - (diff > 0.0 ? sqrt(diff) : 0.0));
            }
            sol[k] = r0[k];
        }
        for (int rec = 0; rec < J.lincs_sv; rec += 2) {   // r0 -> r1, then r1 -> r0
            lincs_expand(J, nc, b, r0, r1, sol);
            if (rec + 1 < J.lincs_sv) lincs_expand(J, nc, b, r1, r0, sol);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k >= nc) continue;
            const int i = J.sc_ij[2 * k], j = J.sc_ij[2 * k + 1];
            const double corr = J.lc_sdiag[k] * sol[k];
            const double si = corr * (i == 0 ? im[0] : (i == 1 ? im[1] : (i == 2 ? im[2] : im[3])));
            const double sj = corr * (j == 0 ? im[0] : (j == 1 ? im[1] : (j == 2 ? im[2] : im[3])));
            const v3 ni = pick4(xn, i);
            put4(xn, i, mk(ni.x + b[k].x * si, ni.y + b[k].y * si, ni.z + b[k].z * si));
            const v3 nj = pick4(xn, j);
            put4(xn, j, mk(nj.x - b[k].x * sj, nj.y - b[k].y * sj, nj.z - b[k].z * sj));
        }
    }
}

// The solvent molecules constrained by SHAKE or LINCS with the template (at
// most 4 atoms and 6 constraints), a lane per group after k_integrate, which
// moved them unconstrained: shake_molecule (its virial into the group's row
// for a pressure-coupled job; an unconverged molecule fails its job) or
// lincs_molecule (no virial: the reference's LINCS has none), then the new
// positions, the constrained atoms' velocities (x - x_old) / dt and the
// group's E_kin terms again, in k_integrate's order.
__global__ void __launch_bounds__(WAVE) k_solvent_constraints(SysPtrs p, int32_t ncg) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncg) return;
    const int32_t jb = p.cg_job[c];
    const JobDev& J = p.jobs[jb];
    if ((J.lincs_sv == 0 && !J.shake_sv) || c - J.cg_off < J.n_solute_cg) return;
    const int32_t step = *p.step;
    if (step >= J.stop) return;   // a frozen job keeps its state
    const int32_t s = p.cg_start[c], n = p.cg_n[c];
    v3 xo[4], xn[4], vn[4];
    double im[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        xo[k] = xn[k] = vn[k] = mk(0.0, 0.0, 0.0);
        im[k] = 0.0;
        if (k < n) {
            xo[k] = ldv(p.x_cur, s + k);
            xn[k] = ldv(p.x_new, s + k);
            vn[k] = ldv(p.v_new, s + k);
            im[k] = p.inv_mass[s + k];
        }
    }
    if (J.shake_sv) {
        double sv[3] = {0.0, 0.0, 0.0};
        if (!shake_molecule(J, n, xo, xn, vn, im, sv)) atomicMin(p.settle_fail + jb, step);
        if (J.baro)
            for (int a = 0; a < 3; ++a) p.cg_cvir[3 * (int64_t)c + a] = sv[a];
    } else {
        unsigned constrained = 0u;
        lincs_molecule(J, xo, xn, im, constrained);
        const double dti = 1.0 / J.dt;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (constrained >> k & 1u)
                vn[k] = mk((xn[k].x - xo[k].x) * dti, (xn[k].y - xo[k].y) * dti, (xn[k].z - xo[k].z) * dti);
    }
    double ek = 0.0, ekn = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= n) continue;
        stv(p.x_new, s + k, xn[k]);
        stv(p.v_new, s + k, vn[k]);
        const v3 vo = ldv(p.v_cur, s + k);
        const double a2 = vn[k].x * vn[k].x + vn[k].y * vn[k].y + vn[k].z * vn[k].z;
        const double b2 = vo.x * vo.x + vo.y * vo.y + vo.z * vo.z;
        ek += 0.5 * p.mass[s + k] * (a2 + b2) / 2.0;   // temperature.rs:76-80, k_integrate's order
        ekn += 0.5 * p.mass[s + k] * a2;
    }
    p.cg_ek[3 * (int64_t)c] = ek;
    p.cg_ek[3 * (int64_t)c + 1] = ekn;
}

// The bath's scale of each coupled job (Berendsen, berendsen_thermostat.rs:66-100,
// or Nose-Hoover, nosehoover_thermostat.rs:116-227) from E_kin_new = 0.5 sum m
// |v|^2 of the velocities entering the step (what the previous
// TemperatureCalculation stored, temperature.rs:76-85).  k_reduce sets it after
// every step; k_therm, block per job, after velocities changed outside (a
// Nose-Hoover chain advances from its committed state either way, so it moves
// once per step).
__global__ void __launch_bounds__(BLOCK) k_therm(SysPtrs p) {
    __shared__ double lds[17];
    const JobDev& J = p.jobs[blockIdx.x];
    if (!J.thermo) {
        if (threadIdx.x == 0) p.therm_scale[blockIdx.x] = 1.0;
        return;
    }
    double s = 0.0;
    for (int32_t a = J.atom_off + threadIdx.x; a < J.atom_off + J.n_atoms; a += blockDim.x) {
        const double* v = p.v_cur + 3 * (int64_t)a;
        s += 0.5 * p.mass[a] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double t[1] = {s};
    block_sum<1>(t, lds);
    if (threadIdx.x == 0) p.therm_scale[blockIdx.x] = bath_scale(p, J, blockIdx.x, lds[16]);
}

// RemoveCOMMotion::apply (remove_com_motion.rs:246-313), first in the md step:
// a block per job, at the job's step 0 (NTICOM) and every |NSCM| steps, on the
// velocities entering the step.  Translation: v -= sum m v / sum m
// (remove_com_motion.rs:90-113).  Rotation (remove_com_motion.rs:129-240): the
// positions, a rectangular box's put into (-L/2, L/2] and taken back half a
// step, about their centre of mass; omega = I^-1 L by the adjugate (zero when
// the determinant is below 1e-12); v -= omega x r.
__global__ void __launch_bounds__(BLOCK) k_com(SysPtrs p) {
    __shared__ double lds[17 * 9];
    const int32_t jb = blockIdx.x;
    const JobDev& J = p.jobs[jb];
    if (J.nticom < 1 && J.nscm == 0) return;
    const int32_t step = *p.step;
    if (step >= J.stop) return;
    const int64_t js = p.job_step0[jb] + step;
    const int64_t skip = J.nscm > 0 ? (int64_t)J.nscm : -(int64_t)J.nscm;
    bool trans, rot;
    if (js == 0) {
        trans = J.nticom >= 1;
        rot = J.nticom >= 2;
    } else if (skip > 0 && js % skip == 0) {
        trans = true;
        rot = J.nscm < 0 && J.box.kind == 0;   // periodic boxes: no periodic rotation removal
    } else {
        return;
    }
    const int32_t a0 = J.atom_off, a1 = J.atom_off + J.n_atoms;
    // the removal changes the step's entering velocities in place (the
    // reference's conf.current().vel), before k_integrate reads them
    double* v = const_cast<double*>(p.v_cur);
    const double* x = p.x_cur;
    if (trans) {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int32_t a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
            const double m = p.mass[a];
            t[0] += m;
            for (int c = 0; c < 3; ++c) t[1 + c] += m * v[3 * (int64_t)a + c];
        }
        block_sum<4>(t, lds);
        const double cv[3] = {lds[65] / lds[64], lds[66] / lds[64], lds[67] / lds[64]};
        __syncthreads();
        for (int32_t a = a0 + threadIdx.x; a < a1; a += blockDim.x)
            for (int c = 0; c < 3; ++c) v[3 * (int64_t)a + c] -= cv[c];
    }
    if (!rot) return;
    const bool wrap = J.box.kind == 1;
    const double dt = J.dt;
    auto wx = [&](int32_t a, int c) {
        const double q = x[3 * (int64_t)a + c];
        return wrap ? q - round(q / J.box.L[c]) * J.box.L[c] : q;
    };
    double t[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int32_t a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
        const double m = p.mass[a];
        t[0] += m;
        for (int c = 0; c < 3; ++c) {
            const double vc = v[3 * (int64_t)a + c];
            t[1 + c] += m * vc;
            t[4 + c] += m * wx(a, c) - 0.5 * m * vc * dt;
        }
    }
    block_sum<7>(t, lds);
    double cv[3], cr[3];
    for (int c = 0; c < 3; ++c) {
        cv[c] = lds[112 + 1 + c] / lds[112];
        cr[c] = lds[112 + 4 + c] / lds[112];
    }
    __syncthreads();
    double u[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // L, I00 I11 I22 I01 I02 I12
    for (int32_t a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
        const double m = p.mass[a];
        double r[3], w[3];
        for (int c = 0; c < 3; ++c) {
            const double vc = v[3 * (int64_t)a + c];
            r[c] = wx(a, c) - 0.5 * dt * vc - cr[c];
            w[c] = vc - cv[c];
        }
        u[0] += m * (r[1] * w[2] - r[2] * w[1]);
        u[1] += m * (r[2] * w[0] - r[0] * w[2]);
        u[2] += m * (r[0] * w[1] - r[1] * w[0]);
        u[3] += m * (r[1] * r[1] + r[2] * r[2]);
        u[4] += m * (r[0] * r[0] + r[2] * r[2]);
        u[5] += m * (r[0] * r[0] + r[1] * r[1]);
        u[6] -= m * r[0] * r[1];
        u[7] -= m * r[0] * r[2];
        u[8] -= m * r[1] * r[2];
    }
    block_sum<9>(u, lds);
    const double* q = lds + 144;
    const double i00 = q[3], i11 = q[4], i22 = q[5], i01 = q[6], i02 = q[7], i12 = q[8];
    const double den = -i02 * i02 * i11 + 2.0 * i01 * i02 * i12 - i00 * i12 * i12 - i01 * i01 * i22 + i00 * i11 * i22;
    const double b00 = -i12 * i12 + i11 * i22, b01 = i02 * i12 - i01 * i22, b02 = -i02 * i11 + i01 * i12;
    const double b11 = -i02 * i02 + i00 * i22, b12 = i01 * i02 - i00 * i12, b22 = -i01 * i01 + i00 * i11;
    double om[3] = {0.0, 0.0, 0.0};
    if (!(den < 1e-12)) {
        om[0] = (b00 * q[0] + b01 * q[1] + b02 * q[2]) / den;
        om[1] = (b01 * q[0] + b11 * q[1] + b12 * q[2]) / den;
        om[2] = (b02 * q[0] + b12 * q[1] + b22 * q[2]) / den;
    }
    for (int32_t a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
        double r[3];
        double* va = v + 3 * (int64_t)a;
        for (int c = 0; c < 3; ++c) r[c] = wx(a, c) - 0.5 * dt * va[c] - cr[c];
        va[0] -= om[1] * r[2] - om[2] * r[1];
        va[1] -= om[2] * r[0] - om[0] * r[2];
        va[2] -= om[0] * r[1] - om[1] * r[0];
    }
}

// prepare_virial's kinetic energy tensor (forcefield.rs:285-296, 1064-1120) of
// each pressure-coupled job, from the velocities entering the step (after the
// centre-of-mass removal): molecular, 0.5 (sum m v)(sum m v)^T / M per
// pressure group, or atomic, 0.5 m v v^T per atom.  A block per job.
__global__ void __launch_bounds__(BLOCK) k_press_ke(SysPtrs p) {
    __shared__ double lds[17 * 9];
    const int32_t jb = blockIdx.x;
    const JobDev& J = p.jobs[jb];
    if (!J.baro || J.vtype == 0) return;
    if (*p.step >= J.stop) return;
    const double* v = p.v_cur;
    double ke[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (J.vtype == 2 && J.n_pg > 0) {
        for (int32_t g = threadIdx.x; g < J.n_pg; g += blockDim.x) {
            const int32_t a0 = p.pg_start[J.pg_off + g], a1 = p.pg_start[J.pg_off + g + 1];
            double mv[3] = {0.0, 0.0, 0.0}, mt = 0.0;
            for (int32_t a = a0; a < a1; ++a) {
                const double m = p.mass[a];
                mt += m;
                for (int c = 0; c < 3; ++c) mv[c] += m * v[3 * (int64_t)a + c];
            }
            if (mt > 0.0)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) ke[3 * a + b] += 0.5 * mv[a] * mv[b] / mt;
        }
    } else {
        for (int32_t i = J.atom_off + threadIdx.x; i < J.atom_off + J.n_atoms; i += blockDim.x) {
            const double m = p.mass[i];
            const double* vi = v + 3 * (int64_t)i;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) ke[3 * a + b] += 0.5 * m * vi[a] * vi[b];
        }
    }
    block_sum<9>(ke, lds);
    if (threadIdx.x < 9) p.press[10 * (int64_t)jb + threadIdx.x] = lds[144 + threadIdx.x];
}

// The pressure and the Berendsen barostat at the end of the md step
// (md.rs:1259-1277), a block per pressure-coupled job, after k_reduce (whose
// last block advanced the step counter, so this step is the counter - 1):
//  * the molecular virial (apply_molecular_virial_correction,
//    forcefield.rs:1128-1200): per pressure group the centre of mass of the
//    chain-gathered positions, then vir -= f_i (x) r_i with r_i the atom's image
//    nearest that centre (the trace: the isotropic pressure reads no more);
//  * the constraints' atomic virial (SETTLE, constraints.rs:978-1010), per
//    group from k_integrate, added after that correction as in the reference;
//  * PressureCalculation (pressure_calculation.rs:40-80): P = tr(KE + vir / 2)
//    * 2 / V / 3 with this step's box;
//  * BerendsenBarostat (berendsen_barostat.rs:30-70): mu = (1 - comp dt / tau_p
//    (P0 - P))^(1/3) scales the new positions and the box (edge lengths and
//    their reciprocals, Rectangular::new) and refreshes the CG-grouped
//    predicate (forcefield.rs:243-283) for the next step.  A box whose cells
//    would fall below half the list cutoff stops the job (box_fail).
__global__ void __launch_bounds__(BLOCK) k_baro(SysPtrs p) {
    __shared__ double lds[17 * 6];
    __shared__ double s_mu;
    const int32_t jb = blockIdx.x;
    // the barostat changes the job's box in place, read by the next step's kernels
    JobDev& J = const_cast<JobDev&>(p.jobs[jb]);
    if (!J.baro) return;
    const int32_t step = *p.step - 1;
    if (step >= J.stop) return;
    const Box box = J.box;
    const double* x = p.x_cur;
    double corr[3] = {0.0, 0.0, 0.0};
    if (J.vtype == 2) {
        for (int32_t g = threadIdx.x; g < J.n_pg; g += blockDim.x) {
            const int32_t a0 = p.pg_start[J.pg_off + g], a1 = p.pg_start[J.pg_off + g + 1];
            d3 prev = ld3(x, a0);
            double com[3] = {0.0, 0.0, 0.0}, mt = 0.0;
            for (int32_t a = a0; a < a1; ++a) {
                const double m = p.mass[a];
                mt += m;
                const d3 r = nim_exact(box, ld3(x, a), prev);
                const d3 gth{r.x + prev.x, r.y + prev.y, r.z + prev.z};
                com[0] += m * gth.x;
                com[1] += m * gth.y;
                com[2] += m * gth.z;
                prev = gth;
            }
            if (mt > 0.0)
                for (int c = 0; c < 3; ++c) com[c] /= mt;
            const d3 cm{com[0], com[1], com[2]};
            for (int32_t a = a0; a < a1; ++a) {
                const d3 r = nim_exact(box, ld3(x, a), cm);
                const double* f = p.f + 3 * (int64_t)a;
                corr[0] += f[0] * r.x;
                corr[1] += f[1] * r.y;
                corr[2] += f[2] * r.z;
            }
        }
    }
    // SETTLE's virial, per group from k_integrate
    double w[6] = {corr[0], corr[1], corr[2], 0.0, 0.0, 0.0};
    for (int32_t c = J.cg_off + threadIdx.x; c < J.cg_off + J.n_cg; c += blockDim.x)
        for (int a = 0; a < 3; ++a) w[3 + a] += p.cg_cvir[3 * (int64_t)c + a];
    block_sum<6>(w, lds);
    if (threadIdx.x == 0) {
        double* pr = p.press + 10 * (int64_t)jb;
        // BoxConfig::volume, the box matrix's determinant c . (a x b) in glam's
        // order (configuration.rs:86-89); the rectangular product equals it
        const double V = box.kind == 2 ? box.c[0] * (box.a[1] * box.b[2] - box.b[1] * box.a[2]) +
                                             box.c[1] * (box.a[2] * box.b[0] - box.b[2] * box.a[0]) +
                                             box.c[2] * (box.a[0] * box.b[1] - box.b[0] * box.a[1])
                                       : box.L[0] * box.L[1] * box.L[2];
        double P = pr[9];   // no virial: the pressure keeps its last value (0 at start)
        if (J.vtype != 0 && !(V < 1e-30)) {
            const double* rec = p.records + ((int64_t)step * gridDim.x + jb) * REC;
            const double factor = 2.0 / V;
            double pt[3];
            for (int a = 0; a < 3; ++a) {
                const double vir = (rec[5 + 4 * a] - lds[96 + a]) + lds[96 + 3 + a];
                pt[a] = (pr[4 * a] + vir * 0.5) * factor;
            }
            P = (pt[0] + pt[1] + pt[2]) / 3.0;
            pr[9] = P;
        }
        const double mu = pow(1.0 - J.comp * J.dt / J.tau_p * (J.p_ref - P), 1.0 / 3.0);
        s_mu = mu;
        if (box.kind == 2) {
            // every box vector scaled (berendsen_barostat.rs:40-48); the diagonal,
            // its inverse and the reduction ratios follow as the box was built
            // (one cell, CG-grouped loops off: nothing else depends on the size)
            for (int a = 0; a < 3; ++a) {
                J.box.a[a] = box.a[a] * mu;
                J.box.b[a] = box.b[a] * mu;
                J.box.c[a] = box.c[a] * mu;
            }
            J.box.L[0] = J.box.a[0];
            J.box.L[1] = J.box.b[1];
            J.box.L[2] = J.box.c[2];
            for (int a = 0; a < 3; ++a) J.box.invL[a] = 1.0 / J.box.L[a];
            J.box.tr[0] = J.box.c[1] / J.box.c[2];
            J.box.tr[1] = J.box.b[0] / J.box.b[1];
            J.box.tr[2] = J.box.c[0] / J.box.c[2];
        } else {
            double hm = 1e300;
            for (int a = 0; a < 3; ++a) {
                const double L = box.L[a] * mu;
                J.box.L[a] = L;
                J.box.invL[a] = 1.0 / L;
                hm = fmin(hm, L * 0.5);
                if (!(L / J.ncell[a] >= J.cell_min)) atomicMin(p.box_fail + jb, step);
            }
            J.use_cg_grouped = hm > J.cgg_limit;
        }
    }
    __syncthreads();
    const double mu = s_mu;
    double* xn = p.x_new;
    for (int64_t k = 3 * (int64_t)J.atom_off + threadIdx.x; k < 3 * (int64_t)(J.atom_off + J.n_atoms); k += blockDim.x)
        xn[k] *= mu;
}

// One md step's update of one original charge group (thread per group): the
// total force (k_assemble), LeapFrogVelocity with the bath's scale and
// LeapFrogPosition (k_leapfrog), SETTLE for rigid waters (k_settle), the E_kin
// terms per group (k_ekin), and - when the next step keeps the pair list - the
// sorted copies xq / xf it reads (k_gather).  A frozen job keeps its state.
template <bool MIXED>
// with_x: add the excluded / 1-4 forces (solute atoms; a solvent atom has none in
// md steps, where k_excl14 leaves its slot unwritten)
__device__ __forceinline__ v3 total_force(const SysPtrs& p, int32_t i, int32_t with_long, bool with_x) {
    int32_t a = p.o2s[i];
    double f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (MIXED) {
            long long t = p.fi_s[3 * a + c] + p.fi_near[3 * a + c];
            if (with_long) t += p.fi_long_s[3 * a + c];
            f[c] = (double)t * FX_INV;
        } else {
            f[c] = p.f_s[3 * a + c];
            if (with_long) f[c] += p.f_long_s[3 * a + c];
        }
        if (with_x) f[c] += p.f_x[3 * i + c];
    }
    return mk(f[0], f[1], f[2]);
}

template <bool MIXED>
__global__ void __launch_bounds__(BLOCK) k_integrate(SysPtrs p, int32_t ncg, int32_t with_long, int32_t write_xq) {
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncg) return;
    const int32_t step = *p.step;   // advanced by the last block of this step's k_reduce
    int32_t jb = p.cg_job[c];
    const JobDev& J = p.jobs[jb];
    const int32_t s = p.cg_start[c], n = p.cg_n[c];
    const bool frozen = step >= J.stop;
    const bool solute_group = c - J.cg_off < J.n_solute_cg;
    const double dt = J.dt;
    const double lam = J.thermo ? p.therm_scale[jb] : 1.0;
    if (J.thermo == 2 && !frozen && c == J.cg_off) {
        // this step applies the chain its scale came from: commit it
        const double* pd = p.nh_pend + (int64_t)NHC_MAX * jb;
        double* zd = p.nh_zeta + (int64_t)NHC_MAX * jb;
        for (int i = 0; i < J.nhc; ++i) zd[i] = pd[i];
    }
    double ek = 0.0, ekn = 0.0;
    auto step_atom = [&](int32_t i, v3& xo, v3& xn, v3& vo, v3& vn) {
        xo = ldv(p.x_cur, i);
        vo = ldv(p.v_cur, i);
        if (frozen) {
            xn = xo;
            vn = vo;
            return;
        }
        v3 f = total_force<MIXED>(p, i, with_long, solute_group);
        if (solute_group && p.bd.fb) {
            // covalent + nonbonded (forcefield.rs:966)
            const double* fb = p.bd.fb + 3 * (int64_t)(J.sol_off + (i - J.atom_off));
            f = mk(fb[0] + f.x, fb[1] + f.y, fb[2] + f.z);
        }
        // a non-finite or absurd force (a diverged job): the job fails at this step
        if (!(fabs(f.x) < FORCE_LIMIT && fabs(f.y) < FORCE_LIMIT && fabs(f.z) < FORCE_LIMIT))
            atomicMin(p.force_fail + jb, step);
        if (J.vtype == 2) stv(p.f, i, f);   // the molecular virial's forces (k_baro)
        double im = p.inv_mass[i];
        if (J.thermo) vn = mk((vo.x + f.x * im * dt) * lam, (vo.y + f.y * im * dt) * lam, (vo.z + f.z * im * dt) * lam);
        else vn = mk(vo.x + f.x * im * dt, vo.y + f.y * im * dt, vo.z + f.z * im * dt);
        xn = mk(xo.x + vn.x * dt, xo.y + vn.y * dt, xo.z + vn.z * dt);
    };
    auto finish_atom = [&](int32_t i, v3 xn, v3 vo, v3 vn, v3 x_first) {
        stv(p.x_new, i, xn);
        stv(p.v_new, i, vn);
        double a2 = vn.x * vn.x + vn.y * vn.y + vn.z * vn.z;
        double b2 = vo.x * vo.x + vo.y * vo.y + vo.z * vo.z;
        ek += 0.5 * p.mass[i] * (a2 + b2) / 2.0;   // temperature.rs:76-80
        ekn += 0.5 * p.mass[i] * a2;
        if (write_xq) {
            int32_t a = p.o2s[i];
            double q = p.q[i];
            p.xq[a] = make_double4(xn.x, xn.y, xn.z, q);
            if (MIXED)
                p.xf[a] = make_float4((float)(xn.x - x_first.x), (float)(xn.y - x_first.y), (float)(xn.z - x_first.z),
                                      (float)q);
        }
    };
    double cv[3] = {0.0, 0.0, 0.0};   // the constraints' virial diagonal (pressure-coupled jobs)
    // (a solvent molecule constrained by SHAKE or LINCS moves unconstrained here;
    // k_solvent_constraints follows)
    if (J.settle && n == 3 && c - J.cg_off >= J.n_solute_cg) {
        v3 xo[3], xn[3], vo[3], vn[3];
        for (int k = 0; k < 3; ++k) step_atom(s + k, xo[k], xn[k], vo[k], vn[k]);
        const v3 xu[3] = {xn[0], xn[1], xn[2]};
        if (!frozen && !settle_one(J, xo, xn, vn)) atomicMin(p.settle_fail + jb, step);
        if (J.baro && !frozen) {
            // constraints.rs:978-1010: constraint forces d m / dt^2 of the displacement,
            // virial += x_old (x) f at the step's starting positions
            const double dti = 1.0 / J.dt, dt2i = dti * dti;
            const double mm[3] = {J.m_o, J.m_h, J.m_h};
            for (int k = 0; k < 3; ++k) {
                cv[0] += xo[k].x * ((xn[k].x - xu[k].x) * mm[k] * dt2i);
                cv[1] += xo[k].y * ((xn[k].y - xu[k].y) * mm[k] * dt2i);
                cv[2] += xo[k].z * ((xn[k].z - xu[k].z) * mm[k] * dt2i);
            }
        }
        for (int k = 0; k < 3; ++k) finish_atom(s + k, xn[k], vo[k], vn[k], xn[0]);
    } else {
        v3 x_first = mk(0.0, 0.0, 0.0);
        for (int32_t k = 0; k < n; ++k) {
            v3 xo, xn, vo, vn;
            step_atom(s + k, xo, xn, vo, vn);
            if (k == 0) x_first = xn;
            finish_atom(s + k, xn, vo, vn, x_first);
        }
    }
    // rf_excluded_interactions for a solvent molecule (excluded.rs:229-262): energy
    // only, -q_i q_j fpepsi crf_2cut3i r^2 for its excluded pairs j > i at the
    // step's starting positions
    double ex = 0.0;
    if (c - J.cg_off >= J.n_solute_cg)
        for (int32_t i = s; i < s + n; ++i) {
            const d3 pi = ld3(p.x_cur, i);
            const double qi = p.q[i];
            for (int32_t k = p.excl_off[i]; k < p.excl_off[i + 1]; ++k) {
                const int32_t j = p.excl[k];
                if (j <= i) continue;
                const d3 r = nim(J.box, pi, ld3(p.x_cur, j));
                const double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
                ex += -qi * p.q[j] * J.fpepsi * J.crf.crf_2cut3i * r2;
            }
        }
    p.cg_ek[3 * (int64_t)c] = ek;
    p.cg_ek[3 * (int64_t)c + 1] = ekn;
    p.cg_ek[3 * (int64_t)c + 2] = ex;
    if (J.baro)
        for (int a = 0; a < 3; ++a) p.cg_cvir[3 * (int64_t)c + a] = cv[a];
}


// One thread's rows of a part, the flags fixed so that the unrolled body has no
// branch and several rows' loads are in flight at once (the additions keep
// their order)
template <bool VIR, bool KIN>
__device__ __forceinline__ void vv_part_rows(const SysPtrs& p, int64_t k0, int64_t k1, double (&v)[PART_W]) {
#pragma unroll 4
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const double* r = p.vv_e + 11 * k;
#pragma unroll
        for (int c = 0; c < (VIR ? 11 : 2); ++c) v[c] += r[c];
        if (p.vv_en) {   // mixed: the near solvent pairs' record
            const double* rn = p.vv_en + 11 * k;
#pragma unroll
            for (int c = 0; c < (VIR ? 11 : 2); ++c) v[c] += rn[c];
        }
        if (KIN) {
            v[1] += p.cg_ek[3 * k + 2];   // solvent excluded RF energy (md steps)
            v[11] += p.cg_ek[3 * k];
            v[12] += p.cg_ek[3 * k + 1];
        }
    }
}

// Per-group solvent-solvent records summed in VV_PARTS contiguous parts per job
// (grid VV_PARTS x njobs), so that the per-job reduction reads few values.
// Per job and part: columns 0-10 from the solvent-solvent records (sorted
// groups), 11-12 the E_kin terms of k_integrate (original groups).
__global__ void __launch_bounds__(BLOCK) k_vv_parts(SysPtrs p, int32_t with_virial, int32_t with_kin) {
    __shared__ double lds[17 * PART_W];
    const JobDev& J = p.jobs[blockIdx.y];
    int64_t n = J.n_cg;
    int64_t k0 = J.cg_off + n * blockIdx.x / VV_PARTS, k1 = J.cg_off + n * (blockIdx.x + 1) / VV_PARTS;
    double v[PART_W] = {0};
    if (with_virial) {
        if (with_kin) vv_part_rows<true, true>(p, k0, k1, v);
        else vv_part_rows<true, false>(p, k0, k1, v);
    } else {
        if (with_kin) vv_part_rows<false, true>(p, k0, k1, v);
        else vv_part_rows<false, false>(p, k0, k1, v);
    }
    block_sum<PART_W>(v, lds);
    if (threadIdx.x < PART_W)
        p.vv_part[((int64_t)blockIdx.y * VV_PARTS + blockIdx.x) * PART_W + threadIdx.x] = lds[16 * PART_W + threadIdx.x];
}

// One block per job: sums the job's partial rows in a fixed order and writes
// the step record (e_lj, e_crf, e_pot, e_kin, dhdl, vir[9]).  Column 11 of a
// row is E_kin in the kinetic slot and dH/dlambda in the others.
// advance: an md step (records row = the device step counter, which the last
// block to finish advances); otherwise a single point (row 0, counter untouched)
__global__ void __launch_bounds__(BLOCK) k_reduce(SysPtrs p, int32_t npad, int32_t with_long,
                                                  int32_t with_kin, int32_t advance, int32_t njobs) {
    constexpr int W = ROW + 2 + 4;   // + dH/dlambda, E_kin_new, the four covalent energies
    __shared__ double lds[17 * W];
    const int32_t step = advance ? *p.step : 0;
    int32_t jb = blockIdx.x;
    const JobDev& J = p.jobs[jb];
    const int32_t stride = p.nrow_blocks;
    double v[W] = {0};
    // solute atoms' general pairs: per-atom records (short every step, long cached)
    for (int32_t w = J.sol_off + threadIdx.x; w < J.sol_off + J.n_solute_atoms; w += blockDim.x) {
        const double* r = p.sol_e + (int64_t)w * ROW;
#pragma unroll
        for (int c = 0; c < 11; ++c) v[c] += r[c];
        v[12] += r[11];
        if (with_long) {
            const double* q = p.sol_e + ((int64_t)p.nsol + w) * ROW;
#pragma unroll
            for (int c = 0; c < 11; ++c) v[c] += q[c];
            v[12] += q[11];
        }
    }
    // solvent-solvent pairs and E_kin: the job's VV_PARTS partial sums
    if (threadIdx.x < VV_PARTS) {
        const double* r = p.vv_part + ((int64_t)jb * VV_PARTS + threadIdx.x) * PART_W;
#pragma unroll
        for (int c = 0; c < 11; ++c) v[c] += r[c];
        if (with_kin) {
            v[11] += r[11];
            v[13] += r[12];
        }
    }
    for (int s = 0; s < NSLOT; ++s) {
        if (s == SLOT_VV || s == SLOT_KIN) continue;
        if (s == SLOT_LONG && !with_long) continue;
        const int32_t per = BLOCK;
        int32_t b0 = J.atom_off / per;
        // md steps write the excluded-pair rows of the solute atoms' blocks only
        int32_t b1 = (J.atom_off + (s == SLOT_EXCL && advance ? J.n_solute_atoms : J.n_atoms) + per - 1) / per;
#pragma unroll 2
        for (int32_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
            const double* r = p.rows + ((int64_t)s * stride + b) * ROW;
#pragma unroll
            for (int c = 0; c < 11; ++c) v[c] += r[c];
            v[12] += r[11];
        }
    }
    // covalent terms: energy per kind, virial with the nonbonded one
    if (p.bd.tev) {
        int32_t t0 = J.bd_off;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t t1 = t0 + J.bd_n[k];
            for (int32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
                const double* ev = p.bd.tev + BD_EV * (int64_t)t;
                v[14 + k] += ev[0];
#pragma unroll
                for (int c = 0; c < 9; ++c) v[2 + c] += ev[1 + c];
            }
            t0 = t1;
        }
    }
    block_sum<W>(v, lds);
    if (threadIdx.x == 0) {
        double* rec = p.records + ((int64_t)step * njobs + jb) * REC;
        const double* t = lds + 16 * W;
        // update_potential_total (configuration.rs:197-207): bond_total carries
        // every covalent term (forcefield.rs:970-971), then LJ, then CRF
        const double eb = ((t[14] + t[15]) + t[16]) + t[17];
        rec[0] = t[0];
        rec[1] = t[1];
        rec[2] = (eb + t[0]) + t[1];
        rec[3] = t[11];
        rec[4] = t[12];
        for (int c = 0; c < 9; ++c) rec[5 + c] = t[2 + c];
        for (int k = 0; k < 4; ++k) rec[14 + k] = t[14 + k];
        // the next step's Berendsen scale from this step's E_kin_new: what the
        // reference reads after TemperatureCalculation stored it
        if (with_kin && J.thermo) p.therm_scale[jb] = bath_scale(p, J, jb, t[13]);
        if (advance) {
            // every block read the counter at its start: the last one to finish
            // advances it for the next step's kernels and rearms the count
            __threadfence();
            if (atomicAdd(p.step_done, 1u) == gridDim.x - 1) {
                *p.step_done = 0u;
                *p.step = step + 1;
                __threadfence();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// test-mode materializer: the reference's four atom-pair lists
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void k_materialize(SysPtrs p, int32_t ncg, unsigned long long* counts, const int64_t* caps,
                              uint32_t* out0, uint32_t* out1, uint32_t* out2, uint32_t* out3) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncg) return;
    int32_t ck = p.s_cg_orig[k];
    const JobDev& J = p.jobs[p.cg_job[ck]];
    int32_t sk = p.s_cg_start[k], nk = p.s_cg_n[k];
    uint32_t* outs[4] = {out0, out1, out2, out3};
    auto emit = [&](int list, int32_t i, int32_t j) {
        unsigned long long slot = atomicAdd(counts + list, 1ull);
        if (FILL && (int64_t)slot < caps[list]) {
            outs[list][2 * slot] = (uint32_t)(i - J.atom_off);
            outs[list][2 * slot + 1] = (uint32_t)(j - J.atom_off);
        }
    };
    if (ck - J.cg_off < J.n_solute_cg) {
        for (int32_t x = 0; x < nk; ++x)
            for (int32_t y = x + 1; y < nk; ++y) {
                int32_t i = p.s2o[sk + x], j = p.s2o[sk + y];
                if (!excluded_or_14(p, i, j)) emit(0, i, j);
            }
    }
    // solvent entries: near ones at the slot's start, the others at its end
    const int64_t vo = p.vv_off[k], vc = p.vv_cap[k], vn = p.vv_near[k];
    for (int64_t x = 0; x < p.vv_cnt[k]; ++x) {
        const uint32_t ent = p.vv_ent[x < vn ? vo + x : vo + vc - 1 - (x - vn)];
        if (!(ent & ENT_CANON)) continue;
        int32_t m = (int32_t)(ent & ENT_IDX);
        emit(2, p.s2o[sk], p.s2o[m]);   // solvent entries hold the partner's first sorted atom
    }
    for (int32_t e = (int32_t)p.gen_off[k]; e < (int32_t)(p.gen_off[k] + p.gen_cnt[k]); ++e) {
        uint32_t ent = p.gen_ent[e];
        if (!(ent & ENT_CANON)) continue;
        int32_t m = (int32_t)(ent & ENT_IDX);
        int32_t sm = p.s_cg_start[m], nm = p.s_cg_n[m];
        bool lng = ent & ENT_LONG, ss = ent & ENT_SS, vvl = ent & ENT_VVLONG;
        int list = vvl ? 3 : (lng ? 1 : 0);
        for (int32_t x = 0; x < nk; ++x)
            for (int32_t y = 0; y < nm; ++y) {
                int32_t i = p.s2o[sk + x], j = p.s2o[sm + y];
                if (ss && !lng && excluded_or_14(p, i, j)) continue;
                emit(list, i, j);
            }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

void launch_cg_ref(const SysPtrs& p, int32_t ncg, int32_t nctl, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_ref, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg, nctl);
}
void launch_sort_cg(const SysPtrs& p, int32_t ncg, int32_t ncell, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg);
    constexpr int32_t WPB = BLOCK / WAVE;
    hipLaunchKernelGGL(k_cell_order, dim3((unsigned)((ncell + WPB - 1) / WPB)), dim3(BLOCK), 0, s, p, ncell);
}
void launch_sorted_atoms(const SysPtrs& p, int32_t ncg, int32_t mixed, hipStream_t s) {
    hipLaunchKernelGGL(k_sorted_atoms, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg, mixed);
}
void launch_caps(const SysPtrs& p, int32_t ncg, int32_t mode, hipStream_t s) {
    if (mode == 0 || mode == 2) hipLaunchKernelGGL(k_caps, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg, mode == 2 ? 1 : 0);
    else hipLaunchKernelGGL(k_caps_from_counts, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg);
}
void launch_nbl(const SysPtrs& p, int32_t ncg, int32_t ncell, bool fill, double near_r2, hipStream_t s) {
    // one block per cell
    dim3 g = xcd_grid(ncell);
    if (fill) hipLaunchKernelGGL(k_nbl<true>, g, dim3(BLOCK), 0, s, p, ncell, near_r2);
    else hipLaunchKernelGGL(k_nbl<false>, g, dim3(BLOCK), 0, s, p, ncell, near_r2);
    unsigned nb = (unsigned)std::min<int64_t>(256, (ncg + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_totals, dim3(nb), dim3(BLOCK), 0, s, p, ncg);
}
void launch_job_pairs(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_job_pairs, dim3(JOB_PAIR_SLICES, njobs), dim3(BLOCK), 0, s, p);
}
void launch_gather(const SysPtrs& p, int32_t npad, int32_t mixed, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, grid_for(npad), dim3(BLOCK), 0, s, p, npad, mixed);
}
// start / stop events (optional) are recorded by the dispatch itself
// (hipExtLaunchKernelGGL), so timing adds no packets between kernels

template <int APS, bool LJ00, int L>
static void launch_vv_l(const SysPtrs& p, int32_t ncg, int32_t virial, int32_t mixed, int32_t neutral,
                        hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
    // waves: VV_CHUNK / (WAVE / L) per chunk of groups
    const int32_t nw = p.nvv_chunk * (VV_CHUNK / (WAVE / L));
    dim3 g = xcd_grid((nw + BLOCK / WAVE - 1) / (BLOCK / WAVE));
    // without timing events: a plain launch (also inside a captured step graph)
#define GAMD_VV_LAUNCH(KER)                                                                      \
    do {                                                                                         \
        if (e0 || e1) hipExtLaunchKernelGGL(KER, g, dim3(BLOCK), 0, s, e0, e1, 0, p, nw);       \
        else hipLaunchKernelGGL(KER, g, dim3(BLOCK), 0, s, p, nw);                               \
    } while (0)
#define GAMD_VV_PERSIST(KER)                                                                               \
    do {                                                                                                   \
        if (e0 || e1) hipExtLaunchKernelGGL(KER, gp, dim3(BLOCK), 0, s, e0, e1, 0, p, nw, per_xcd);       \
        else hipLaunchKernelGGL(KER, gp, dim3(BLOCK), 0, s, p, nw, per_xcd);                               \
    } while (0)
    if (mixed && p.vv_persist > 0) {
        // the persistent form: at most p.vv_persist resident blocks, a multiple of
        // the XCD count, each XCD's share of logical waves as k_vv_mx splits them
        const int32_t perb = (int32_t)(g.x / NXCD);
        const int32_t per_xcd = perb * (BLOCK / WAVE);
        const dim3 gp((unsigned)(std::max(1, std::min(perb, p.vv_persist / NXCD)) * NXCD));
        if (neutral) {
            if (virial) GAMD_VV_PERSIST((k_vv_mx_persist<APS, LJ00, true, L, true>));
            else GAMD_VV_PERSIST((k_vv_mx_persist<APS, LJ00, false, L, true>));
        } else {
            if (virial) GAMD_VV_PERSIST((k_vv_mx_persist<APS, LJ00, true, L, false>));
            else GAMD_VV_PERSIST((k_vv_mx_persist<APS, LJ00, false, L, false>));
        }
    } else if (mixed && neutral) {
        if (virial) GAMD_VV_LAUNCH((k_vv_mx<APS, LJ00, true, L, true>));
        else GAMD_VV_LAUNCH((k_vv_mx<APS, LJ00, false, L, true>));
    } else if (mixed) {
        if (virial) GAMD_VV_LAUNCH((k_vv_mx<APS, LJ00, true, L, false>));
        else GAMD_VV_LAUNCH((k_vv_mx<APS, LJ00, false, L, false>));
    } else {
        if (virial) GAMD_VV_LAUNCH((k_vv<APS, LJ00, true, L>));
        else GAMD_VV_LAUNCH((k_vv<APS, LJ00, false, L>));
    }
#undef GAMD_VV_LAUNCH
#undef GAMD_VV_PERSIST
}
template <int APS, bool LJ00>
static void launch_vv_aps(const SysPtrs& p, int32_t ncg, int32_t virial, int32_t mixed, int32_t neutral,
                          int32_t lanes, hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
    if (lanes == 4) launch_vv_l<APS, LJ00, 4>(p, ncg, virial, mixed, neutral, e0, e1, s);
    else if (lanes == 16) launch_vv_l<APS, LJ00, 16>(p, ncg, virial, mixed, neutral, e0, e1, s);
    else launch_vv_l<APS, LJ00, 8>(p, ncg, virial, mixed, neutral, e0, e1, s);
}
template <int APS, bool LJ00>
static void launch_vv_near_aps(const SysPtrs& p, int32_t virial, hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
    const int32_t nw = p.nvv_chunk * (VV_CHUNK / (WAVE / 4));   // 4 lanes per molecule
    dim3 g = xcd_grid((nw + BLOCK / WAVE - 1) / (BLOCK / WAVE));
    if (e0 || e1) {
        if (virial) hipExtLaunchKernelGGL((k_vv_nr<APS, LJ00, true, 4>), g, dim3(BLOCK), 0, s, e0, e1, 0, p, nw);
        else hipExtLaunchKernelGGL((k_vv_nr<APS, LJ00, false, 4>), g, dim3(BLOCK), 0, s, e0, e1, 0, p, nw);
    } else {
        if (virial) hipLaunchKernelGGL((k_vv_nr<APS, LJ00, true, 4>), g, dim3(BLOCK), 0, s, p, nw);
        else hipLaunchKernelGGL((k_vv_nr<APS, LJ00, false, 4>), g, dim3(BLOCK), 0, s, p, nw);
    }
}
void launch_vv_near(const SysPtrs& p, int32_t aps, int32_t form, int32_t virial, hipStream_t s, hipEvent_t e0,
                    hipEvent_t e1) {
    switch (aps * 2 + ((form & VV_LJ00) ? 1 : 0)) {
        case 2: launch_vv_near_aps<1, false>(p, virial, e0, e1, s); break;
        case 3: launch_vv_near_aps<1, true>(p, virial, e0, e1, s); break;
        case 4: launch_vv_near_aps<2, false>(p, virial, e0, e1, s); break;
        case 5: launch_vv_near_aps<2, true>(p, virial, e0, e1, s); break;
        case 8: launch_vv_near_aps<4, false>(p, virial, e0, e1, s); break;
        case 9: launch_vv_near_aps<4, true>(p, virial, e0, e1, s); break;
        case 7: launch_vv_near_aps<3, true>(p, virial, e0, e1, s); break;
        default: launch_vv_near_aps<3, false>(p, virial, e0, e1, s); break;
    }
}
void launch_vv(const SysPtrs& p, int32_t ncg, int32_t aps, int32_t form, int32_t virial, int32_t mixed,
               int32_t lanes, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const int32_t neutral = (form & VV_NEUTRAL) ? 1 : 0;
    switch (aps * 2 + ((form & VV_LJ00) ? 1 : 0)) {
        case 2: launch_vv_aps<1, false>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 3: launch_vv_aps<1, true>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 4: launch_vv_aps<2, false>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 5: launch_vv_aps<2, true>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 8: launch_vv_aps<4, false>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 9: launch_vv_aps<4, true>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        case 7: launch_vv_aps<3, true>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
        default: launch_vv_aps<3, false>(p, ncg, virial, mixed, neutral, lanes, e0, e1, s); break;
    }
}
void launch_gen(const SysPtrs& p, int32_t npad, bool lng, int32_t virial, int32_t mixed, int32_t ngsol,
                hipStream_t s, int32_t part) {
    const bool solv = part != GEN_SOLUTE, solu = part != GEN_SOLVENT;
    // solvent atoms a lane each, then solute atoms a wave each (after: the long
    // pass of k_gen writes zeros for the solute atoms)
    const int32_t nsol = p.nsol;
    dim3 gs((unsigned)nsol);   // a block per solute atom
    if (mixed) {
        if (lng) {
            hipLaunchKernelGGL((k_gen_mx<true, true>), grid_for(npad), dim3(BLOCK), 0, s, p, npad);
            if (nsol) hipLaunchKernelGGL((k_gen_sol<true, true, true, BLOCK>), gs, dim3(BLOCK), 0, s, p, nsol, 1);
        } else {
            // the short pass of the solvent atoms: only the listed groups
            // two lanes per atom: 46.6 us per step on the 64-job FEP mix, against
            // 55.4 (one lane), 53.4 (four) and 132.5 (sixteen)
            constexpr int32_t GPB = BLOCK / (4 * 2);   // groups per block
            if (solv && ngsol > 0)
                hipLaunchKernelGGL(k_gen_sv<2>, dim3((unsigned)((ngsol + GPB - 1) / GPB)), dim3(BLOCK), 0, s, p, ngsol);
            if (solu && nsol && virial)
                hipLaunchKernelGGL((k_gen_sol<false, true, true, BLOCK>), gs, dim3(BLOCK), 0, s, p, nsol, 1);
            else if (solu && nsol)
                hipLaunchKernelGGL((k_gen_sol<false, true, false, BLOCK>), gs, dim3(BLOCK), 0, s, p, nsol, 0);
        }
        return;
    }
    if (lng) {
        hipLaunchKernelGGL(k_gen<true>, grid_for(npad), dim3(BLOCK), 0, s, p, npad, virial);
        if (nsol) hipLaunchKernelGGL((k_gen_sol<true, false, true, BLOCK>), gs, dim3(BLOCK), 0, s, p, nsol, 1);
    } else {
        hipLaunchKernelGGL(k_gen<false>, grid_for(npad), dim3(BLOCK), 0, s, p, npad, virial);
        if (nsol) hipLaunchKernelGGL((k_gen_sol<false, false, false, BLOCK>), gs, dim3(BLOCK), 0, s, p, nsol, virial);
    }
}
void launch_excl14(const SysPtrs& p, int32_t npad, int32_t solv_energy, const int32_t* blocks, int32_t nblocks,
                   hipStream_t s) {
    if (!blocks) hipLaunchKernelGGL(k_excl14, grid_for(npad), dim3(BLOCK), 0, s, p, npad, solv_energy, blocks);
    else if (nblocks > 0)
        hipLaunchKernelGGL(k_excl14, dim3((unsigned)nblocks), dim3(BLOCK), 0, s, p, npad, solv_energy, blocks);
}
void launch_assemble(const SysPtrs& p, int32_t npad, int32_t with_long, int32_t mixed,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_assemble, grid_for(npad), dim3(BLOCK), 0, s, p, npad, with_long, mixed);
}
void launch_integrate(const SysPtrs& p, int32_t ncg, int32_t with_long, int32_t write_xq,
                      int32_t mixed, hipStream_t s) {
    // one wave per block: a lane per group gives few waves (32,768 waters: 512),
    // spread over all CUs this way rather than 128 blocks on half of them
    const dim3 g((unsigned)((ncg + WAVE - 1) / WAVE));
    if (mixed) hipLaunchKernelGGL(k_integrate<true>, g, dim3(WAVE), 0, s, p, ncg, with_long, write_xq);
    else hipLaunchKernelGGL(k_integrate<false>, g, dim3(WAVE), 0, s, p, ncg, with_long, write_xq);
}
void launch_press_ke(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_press_ke, dim3(njobs), dim3(BLOCK), 0, s, p);
}
void launch_shake_solute(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_shake_solute, dim3(njobs), dim3(BLOCK), 0, s, p);
}
void launch_solvent_constraints(const SysPtrs& p, int32_t ncg, hipStream_t s) {
    hipLaunchKernelGGL(k_solvent_constraints, dim3((unsigned)((ncg + WAVE - 1) / WAVE)), dim3(WAVE), 0, s, p, ncg);
}
void launch_baro(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_baro, dim3(njobs), dim3(BLOCK), 0, s, p);
}
void launch_com(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_com, dim3(njobs), dim3(BLOCK), 0, s, p);
}
void launch_therm(const SysPtrs& p, int32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_therm, dim3(njobs), dim3(BLOCK), 0, s, p);
}
void launch_reduce(const SysPtrs& p, int32_t npad, int32_t with_long, int32_t with_kin, int32_t advance,
                   int32_t njobs, hipStream_t s) {
    // one wave per part: the parts of a small job hold few groups
    hipLaunchKernelGGL(k_vv_parts, dim3(VV_PARTS, njobs), dim3(WAVE), 0, s, p, p.vv_virial, with_kin);
    hipLaunchKernelGGL(k_reduce, dim3(njobs), dim3(BLOCK), 0, s, p, npad, with_long, with_kin, advance, njobs);
}
void launch_materialize(const SysPtrs& p, int32_t ncg, bool fill, unsigned long long* counts, const int64_t* caps,
                        uint32_t* o0, uint32_t* o1, uint32_t* o2, uint32_t* o3, hipStream_t s) {
    if (fill) hipLaunchKernelGGL(k_materialize<true>, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg, counts, caps, o0, o1, o2, o3);
    else hipLaunchKernelGGL(k_materialize<false>, grid_for(ncg), dim3(BLOCK), 0, s, p, ncg, counts, caps, o0, o1, o2, o3);
}

}  // namespace gamd
