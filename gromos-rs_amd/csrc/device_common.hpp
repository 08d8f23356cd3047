// Device-side building blocks shared by the gfx950 kernels.
//
// The FP64 helpers restate the reference arithmetic operation by operation
// (the library is compiled with -ffp-contract=off, so nothing is fused):
//   nearest image        crates/gromos-core/src/math.rs:58-68
//   pair kernel          crates/gromos-forces/src/nonbonded/innerloops.rs:22-55
//   soft-core kernel     crates/gromos-forces/src/nonbonded/perturbed.rs:124-266
// HIP's double division and sqrt are the correctly rounded IEEE operations and
// round() rounds half away from zero like f64::round, so each per-pair value
// equals the reference's bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gamd {

constexpr int WAVE = 64;

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ d3 ld3(const double* p, int64_t i) {
    return d3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}

// Boundary: kind 0 = vacuum, 1 = rectangular, 2 = triclinic.  invL = 1.0 / L
// computed once on the host (DVec3::recip, math.rs:49-55).  Triclinic: the
// lower-triangular box vectors a, b, c (math.rs:75-92); L holds the diagonal.
struct Box {
    double L[3];
    double invL[3];
    int32_t kind;
    int32_t pad;
    double a[3], b[3], c[3];
    double tr[3];   // triclinic: c[1] / c[2], b[0] / b[1], c[0] / c[2]
};

// Triclinic::nearest_image (math.rs:105-141): z, then y, then x reduction loops
// in the reference's operation order.  Each loop is bounded (a lattice vector
// per pass: positions a few boxes apart need a few passes) so that non-finite
// coordinates cannot keep a lane in it.
constexpr int TRI_MAX_PASSES = 1024;
__device__ __forceinline__ d3 nim_tri(const Box& B, d3 n) {
    const double* a = B.a;
    const double* b = B.b;
    const double* c = B.c;
    int k = 0;
    while (n.z >= 0.5 * c[2] && k++ < TRI_MAX_PASSES) { n.x -= c[0]; n.y -= c[1]; n.z -= c[2]; }
    k = 0;
    while (n.z <= -0.5 * c[2] && k++ < TRI_MAX_PASSES) { n.x += c[0]; n.y += c[1]; n.z += c[2]; }
    double ny = n.y - (c[1] / c[2]) * n.z;
    k = 0;
    while (ny >= 0.5 * b[1] && k++ < TRI_MAX_PASSES) {
        n.x -= b[0]; n.y -= b[1]; n.z -= b[2];
        ny = n.y - (c[1] / c[2]) * n.z;
    }
    k = 0;
    while (ny <= -0.5 * b[1] && k++ < TRI_MAX_PASSES) {
        n.x += b[0]; n.y += b[1]; n.z += b[2];
        ny = n.y - (c[1] / c[2]) * n.z;
    }
    double nx = n.x - b[0] * (n.y - c[1] * n.z / c[2]) / b[1] - c[0] * n.z / c[2];
    k = 0;
    while (nx >= 0.5 * a[0] && k++ < TRI_MAX_PASSES) {
        n.x -= a[0]; n.y -= a[1]; n.z -= a[2];
        nx = n.x - b[0] * (n.y - c[1] * n.z / c[2]) / b[1] - c[0] * n.z / c[2];
    }
    k = 0;
    while (nx <= -0.5 * a[0] && k++ < TRI_MAX_PASSES) {
        n.x += a[0]; n.y += a[1]; n.z += a[2];
        nx = n.x - b[0] * (n.y - c[1] * n.z / c[2]) / b[1] - c[0] * n.z / c[2];
    }
    return n;
}

// The same reduction in closed form for the force kernels: the number of
// lattice vectors each loop of nim_tri takes is the rounded ratio of the
// reduced coordinate to the diagonal entry (a[1] = a[2] = b[2] = 0, so the z, y
// and x steps are independent in that order).  In exact arithmetic the vector
// is nim_tri's (ties at exact halves aside); in floating point it differs from
// the loops' repeated subtraction by an ulp or so, with far fewer registers
// and no data-dependent loops.  The pair search keeps the loops (its pair set
// is the reference's to the bit).
__device__ __forceinline__ d3 nim_tri_fast(const Box& B, d3 n) {
    const double kz = rint(n.z * B.invL[2]);
    n.x -= kz * B.c[0];
    n.y -= kz * B.c[1];
    n.z -= kz * B.c[2];
    const double ky = rint((n.y - B.tr[0] * n.z) * B.invL[1]);
    n.x -= ky * B.b[0];
    n.y -= ky * B.b[1];
    n.z -= ky * B.b[2];
    const double kx = rint((n.x - B.tr[1] * (n.y - B.tr[0] * n.z) - B.tr[2] * n.z) * B.invL[0]);
    n.x -= kx * B.a[0];
    n.y -= kx * B.a[1];
    n.z -= kx * B.a[2];
    return n;
}

__device__ __forceinline__ d3 nim(const Box& b, d3 ri, d3 rj) {
    d3 r{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    if (b.kind == 1) {
        r.x = r.x - b.L[0] * round(r.x * b.invL[0]);
        r.y = r.y - b.L[1] * round(r.y * b.invL[1]);
        r.z = r.z - b.L[2] * round(r.z * b.invL[2]);
    } else if (b.kind == 2) {
        r = nim_tri_fast(b, r);
    }
    return r;
}

// nearest image with the reference's triclinic loops (the pair search)
__device__ __forceinline__ d3 nim_exact(const Box& b, d3 ri, d3 rj) {
    if (b.kind == 2) return nim_tri(b, d3{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z});
    return nim(b, ri, rj);
}

struct Crf {
    double crf_cut, crf_2cut3i, crf_cut3i, cutoff_sq;
};

// lj_crf_interaction: returns f/r, e_lj, e_crf; zeros when r2 < 1e-10.
__device__ __forceinline__ void lj_crf64(double x, double y, double z, double c6, double c12,
                                         double qp, const Crf& crf, double& f, double& elj,
                                         double& ecrf) {
    double r2 = x * x + y * y + z * z;
    if (r2 < 1e-10) {
        f = 0.0;
        elj = 0.0;
        ecrf = 0.0;
        return;
    }
    double inv_r2 = 1.0 / r2;
    double inv_r6 = inv_r2 * inv_r2 * inv_r2;
    elj = (c12 * inv_r6 - c6) * inv_r6;
    double flj = (12.0 * c12 * inv_r6 - 6.0 * c6) * inv_r6 * inv_r2;
    double inv_r = sqrt(inv_r2);
    ecrf = qp * (inv_r - crf.crf_2cut3i * r2 - crf.crf_cut);
    double fcrf = qp * (inv_r * inv_r2 + crf.crf_cut3i);
    f = flj + fcrf;
}

// PerturbedLambdaParams field order (perturbed.rs:11-50)
struct Lambda {
    double a_lj_lambda_n, b_lj_lambda_n, a_crf_lambda_n, b_crf_lambda_n;
    double a_lj_lambda_n_1, b_lj_lambda_n_1, a_crf_lambda_n_1, b_crf_lambda_n_1;
    double a_ljs_lambda, b_ljs_lambda, a_ljs_lambda2, b_ljs_lambda2;
    double a_crfs_lambda, b_crfs_lambda, a_crfs_lambda2, b_crfs_lambda2;
    int32_t lambda_exp;
    int32_t pad;
};

// Reaction-field constants of the soft-core pair that depend only on the pair's
// alpha_crf and lambda (perturbed.rs:160-190): hoisted where one perturbed atom
// meets many partners (same expressions, same rounding).
struct SoftRF {
    double a_crf_2cut3i, b_crf_2cut3i, a_crf_cut3i, b_crf_cut3i, a_crf_pert, b_crf_pert;
};
__device__ __forceinline__ SoftRF soft_rf(double alpha_crf, const Lambda& lp, const Crf& crf) {
    SoftRF c;
    double cut2 = crf.cutoff_sq;
    double a_cut2soft = cut2 + alpha_crf * lp.b_crfs_lambda2;
    double b_cut2soft = cut2 + alpha_crf * lp.a_crfs_lambda2;
    double a_cut2soft3 = a_cut2soft * a_cut2soft * a_cut2soft;
    double b_cut2soft3 = b_cut2soft * b_cut2soft * b_cut2soft;
    c.a_crf_2cut3i = crf.crf_2cut3i / sqrt(a_cut2soft3);
    c.b_crf_2cut3i = crf.crf_2cut3i / sqrt(b_cut2soft3);
    c.a_crf_cut3i = 2.0 * c.a_crf_2cut3i;
    c.b_crf_cut3i = 2.0 * c.b_crf_2cut3i;
    c.a_crf_pert = 3.0 * c.a_crf_2cut3i / a_cut2soft;
    c.b_crf_pert = 3.0 * c.b_crf_2cut3i / b_cut2soft;
    return c;
}

// perturbed_lj_crf_interaction: o = {f, e_lj, e_crf, de_lj, de_crf}, with the
// pair's soft_rf constants
__device__ __forceinline__ void pert_lj_crf64_rf(double x, double y, double z, double a_c6,
                                                 double a_c12, double b_c6, double b_c12,
                                                 double a_q, double b_q, double alpha_lj,
                                                 double alpha_crf, const Lambda& lp,
                                                 const Crf& crf, const SoftRF& sc, double fpepsi, double o[5]) {
    double r2 = x * x + y * y + z * z;
    if (r2 < 1e-10) {
        o[0] = o[1] = o[2] = o[3] = o[4] = 0.0;
        return;
    }
    double r6 = r2 * r2 * r2;
    double r4 = r2 * r2;
    double a_c126 = fabs(a_c6) > 1e-10 ? a_c12 / a_c6 : 0.0;
    double b_c126 = fabs(b_c6) > 1e-10 ? b_c12 / b_c6 : 0.0;
    double a_dist2soft = r2 + alpha_crf * lp.b_crfs_lambda2;
    double a_distisoft = 1.0 / sqrt(a_dist2soft);
    double a_dist3isoft = a_distisoft / a_dist2soft;
    double a_dist6soft = r6 + alpha_lj * lp.b_ljs_lambda2 * a_c126;
    double a_dist6isoft = 1.0 / a_dist6soft;
    double b_dist2soft = r2 + alpha_crf * lp.a_crfs_lambda2;
    double b_distisoft = 1.0 / sqrt(b_dist2soft);
    double b_dist3isoft = b_distisoft / b_dist2soft;
    double b_dist6soft = r6 + alpha_lj * lp.a_ljs_lambda2 * b_c126;
    double b_dist6isoft = 1.0 / b_dist6soft;
    const double a_crf_2cut3i = sc.a_crf_2cut3i, b_crf_2cut3i = sc.b_crf_2cut3i;
    const double a_crf_cut3i = sc.a_crf_cut3i, b_crf_cut3i = sc.b_crf_cut3i;
    const double a_crf_pert = sc.a_crf_pert, b_crf_pert = sc.b_crf_pert;
    double force = (lp.a_crf_lambda_n * a_q * (a_dist3isoft + a_crf_cut3i) +
                    lp.b_crf_lambda_n * b_q * (b_dist3isoft + b_crf_cut3i)) *
                   fpepsi;
    force += -6.0 *
             (lp.a_lj_lambda_n * a_c6 * a_dist6isoft * a_dist6isoft +
              lp.b_lj_lambda_n * b_c6 * b_dist6isoft * b_dist6isoft) *
             r4;
    force += 12.0 *
             (lp.a_lj_lambda_n * a_c12 * a_dist6isoft * a_dist6isoft * a_dist6isoft +
              lp.b_lj_lambda_n * b_c12 * b_dist6isoft * b_dist6isoft * b_dist6isoft) *
             r4;
    double a_e_lj = (a_c12 * a_dist6isoft - a_c6) * a_dist6isoft;
    double b_e_lj = (b_c12 * b_dist6isoft - b_c6) * b_dist6isoft;
    double a_e_crf = a_q * (a_distisoft - a_crf_2cut3i * r2 - crf.crf_cut);
    double b_e_crf = b_q * (b_distisoft - b_crf_2cut3i * r2 - crf.crf_cut);
    double e_lj = lp.a_lj_lambda_n * a_e_lj + lp.b_lj_lambda_n * b_e_lj;
    double e_crf = (lp.a_crf_lambda_n * a_e_crf + lp.b_crf_lambda_n * b_e_crf) * fpepsi;
    double de_lj = -2.0 * alpha_lj *
                       (lp.a_lj_lambda_n * lp.b_ljs_lambda * a_c126 * a_dist6isoft *
                            a_dist6isoft * (2.0 * a_c12 * a_dist6isoft - a_c6) -
                        lp.b_lj_lambda_n * lp.a_ljs_lambda * b_c126 * b_dist6isoft *
                            b_dist6isoft * (2.0 * b_c12 * b_dist6isoft - b_c6)) +
                   (double)lp.lambda_exp *
                       (lp.b_lj_lambda_n_1 * b_e_lj - lp.a_lj_lambda_n_1 * a_e_lj);
    double de_crf = -(lp.a_crf_lambda_n * a_q * lp.b_crfs_lambda * (a_dist3isoft - a_crf_pert * r2) -
                      lp.b_crf_lambda_n * b_q * lp.a_crfs_lambda * (b_dist3isoft - b_crf_pert * r2)) *
                        fpepsi * alpha_crf +
                    (double)lp.lambda_exp *
                        (lp.b_crf_lambda_n_1 * b_e_crf - lp.a_crf_lambda_n_1 * a_e_crf) * fpepsi;
    o[0] = force;
    o[1] = e_lj;
    o[2] = e_crf;
    o[3] = de_lj;
    o[4] = de_crf;
}

__device__ __forceinline__ void pert_lj_crf64(double x, double y, double z, double a_c6,
                                              double a_c12, double b_c6, double b_c12,
                                              double a_q, double b_q, double alpha_lj,
                                              double alpha_crf, const Lambda& lp,
                                              const Crf& crf, double fpepsi, double o[5]) {
    pert_lj_crf64_rf(x, y, z, a_c6, a_c12, b_c6, b_c12, a_q, b_q, alpha_lj, alpha_crf, lp, crf,
                     soft_rf(alpha_crf, lp, crf), fpepsi, o);
}

// One pair's contribution to a perturbed correction pass: the scalar force
// delta along r (force delta = r * f), the energy deltas and dH/dlambda.
struct PertDelta {
    double f, de_lj, de_crf, dhdl;
};

// perturbed_self_energy_correction, one perturbed atom (perturbed.rs:406-422):
// the RF self term of the soft-core kernel at r = 0 minus the state-A one.
__device__ __forceinline__ void pert_self64(double a_q, double b_q, const Lambda& lp, const Crf& crf,
                                            double fpepsi, double& de, double& dhdl) {
    double qa2 = a_q * a_q, qb2 = b_q * b_q;
    double e_rf = (-(lp.a_crf_lambda_n * qa2) - lp.b_crf_lambda_n * qb2 + qa2) * crf.crf_cut * fpepsi;
    double de_rf = (double)lp.lambda_exp * (lp.a_crf_lambda_n_1 * qa2 - lp.b_crf_lambda_n_1 * qb2) * crf.crf_cut *
                   fpepsi;
    de = 0.5 * e_rf;
    dhdl = 0.5 * de_rf;
}

// perturbed_excluded_correction, one excluded pair at squared distance r2
// (perturbed.rs:445-510): soft RF cutoff scaled by sqrt((rc^2 / cut2soft)^3),
// no 1/r term; a_q, b_q are the pair's charge products, acrf its alpha_crf.
__device__ __forceinline__ PertDelta pert_excl64(double r2, double a_q, double b_q, double acrf, const Lambda& lp,
                                                 const Crf& crf, double fpepsi) {
    double cut2 = crf.cutoff_sq;
    double a_cut2soft = cut2 + acrf * lp.b_crfs_lambda2;
    double b_cut2soft = cut2 + acrf * lp.a_crfs_lambda2;
    double ra = cut2 / a_cut2soft, rb = cut2 / b_cut2soft;
    double a_crf_2cut3i = crf.crf_2cut3i * sqrt(ra * (ra * ra));
    double b_crf_2cut3i = crf.crf_2cut3i * sqrt(rb * (rb * rb));
    double a_crf_cut3i = 2.0 * a_crf_2cut3i, b_crf_cut3i = 2.0 * b_crf_2cut3i;
    double a_crf_pert = 3.0 * a_crf_2cut3i / a_cut2soft, b_crf_pert = 3.0 * b_crf_2cut3i / b_cut2soft;
    double a_e_crf = a_q * (-a_crf_2cut3i * r2 - crf.crf_cut);
    double b_e_crf = b_q * (-b_crf_2cut3i * r2 - crf.crf_cut);
    double e_crf_p = (lp.a_crf_lambda_n * a_e_crf + lp.b_crf_lambda_n * b_e_crf) * fpepsi;
    double f_p = (lp.a_crf_lambda_n * a_q * a_crf_cut3i + lp.b_crf_lambda_n * b_q * b_crf_cut3i) * fpepsi;
    double e_crf_a = a_q * fpepsi * (-crf.crf_2cut3i * r2 - crf.crf_cut);
    double f_a = a_q * fpepsi * crf.crf_cut3i;
    double n = (double)lp.lambda_exp;
    double dcrf = ((lp.a_crf_lambda_n * a_q * lp.b_crfs_lambda * a_crf_pert -
                    lp.b_crf_lambda_n * b_q * lp.a_crfs_lambda * b_crf_pert) * r2 * acrf +
                   n * (lp.b_crf_lambda_n_1 * b_e_crf - lp.a_crf_lambda_n_1 * a_e_crf)) * fpepsi;
    return PertDelta{f_p - f_a, 0.0, e_crf_p - e_crf_a, dcrf};
}

// perturbed_one_four_correction, one 1-4 pair with r2 >= 1e-10
// (perturbed.rs:565-604): soft-core kernel on the cs6/cs12 entries of both
// states minus the state-A 1-4 term (coulomb scaling 1).
__device__ __forceinline__ PertDelta pert_14_64(double x, double y, double z, double r2, const double* la,
                                                const double* lb, double a_q, double b_q, double alj, double acrf,
                                                const Lambda& lp, const Crf& crf, double fpepsi) {
    double inv_r2 = 1.0 / r2;
    double inv_r6 = inv_r2 * inv_r2 * inv_r2;
    double inv_r = sqrt(inv_r2);
    double e_lj_a = (la[3] * inv_r6 - la[2]) * inv_r6;
    double f_lj_a = (12.0 * la[3] * inv_r6 - 6.0 * la[2]) * inv_r6 * inv_r2;
    double e_crf_a = a_q * fpepsi * (inv_r - crf.crf_2cut3i * r2 - crf.crf_cut);
    double f_crf_a = a_q * fpepsi * (inv_r * inv_r2 + crf.crf_cut3i);
    double o[5];
    pert_lj_crf64(x, y, z, la[2], la[3], lb[2], lb[3], a_q, b_q, alj, acrf, lp, crf, fpepsi, o);
    return PertDelta{o[0] - (f_lj_a + f_crf_a), o[1] - e_lj_a, o[2] - e_crf_a, o[3] + o[4]};
}

// perturbed_atom_pair_correction, one PERTATOMPAIR entry with r2 >= 1e-10
// (perturbed.rs:654-706).  lj = the atoms' regular LJ entry, qp = q_i q_j
// fpepsi, at / bt = interaction type in A / B (0 = c6/c12, 1 = cs6/cs12,
// -1 = absent in B).  The regular 1-4 term is the state the pair leaves.
__device__ __forceinline__ PertDelta pert_atom_pair64(double r2, const double* lj, double qp, int32_t at, int32_t bt,
                                                      const Lambda& lp, const Crf& crf) {
    double inv_r2 = 1.0 / r2;
    double inv_r6 = inv_r2 * inv_r2 * inv_r2;
    double inv_r = sqrt(inv_r2);
    double e_lj_reg = (lj[3] * inv_r6 - lj[2]) * inv_r6;
    double f_lj_reg = (12.0 * lj[3] * inv_r6 - 6.0 * lj[2]) * inv_r6 * inv_r2;
    double e_crf_reg = qp * (inv_r - crf.crf_2cut3i * r2 - crf.crf_cut);
    double f_crf_reg = qp * (inv_r * inv_r2 + crf.crf_cut3i);
    double c6a = at == 0 ? lj[0] : lj[2], c12a = at == 0 ? lj[1] : lj[3];
    double e_lj_a = (c12a * inv_r6 - c6a) * inv_r6;
    double f_lj_a = (12.0 * c12a * inv_r6 - 6.0 * c6a) * inv_r6 * inv_r2;
    double e_lj_b = 0.0, f_lj_b = 0.0, e_crf_b = 0.0, f_crf_b = 0.0;
    if (bt >= 0) {
        double c6b = bt == 0 ? lj[0] : lj[2], c12b = bt == 0 ? lj[1] : lj[3];
        e_lj_b = (c12b * inv_r6 - c6b) * inv_r6;
        f_lj_b = (12.0 * c12b * inv_r6 - 6.0 * c6b) * inv_r6 * inv_r2;
        e_crf_b = e_crf_reg;
        f_crf_b = f_crf_reg;
    }
    double n = (double)lp.lambda_exp;
    double e_lj_p = lp.a_lj_lambda_n * e_lj_a + lp.b_lj_lambda_n * e_lj_b;
    double e_crf_p = lp.a_crf_lambda_n * e_crf_reg + lp.b_crf_lambda_n * e_crf_b;
    double f_lj_p = lp.a_lj_lambda_n * f_lj_a + lp.b_lj_lambda_n * f_lj_b;
    double f_crf_p = lp.a_crf_lambda_n * f_crf_reg + lp.b_crf_lambda_n * f_crf_b;
    double dlj = n * (lp.b_lj_lambda_n_1 * e_lj_b - lp.a_lj_lambda_n_1 * e_lj_a);
    double dcrf = n * (lp.b_crf_lambda_n_1 * e_crf_b - lp.a_crf_lambda_n_1 * e_crf_reg);
    return PertDelta{(f_lj_p - f_lj_reg) + (f_crf_p - f_crf_reg), e_lj_p - e_lj_reg, e_crf_p - e_crf_reg,
                     dlj + dcrf};
}

// ---- wave / block reductions (deterministic order) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Sums NV values over the block into lds[0..NV); every thread must call.
// blockDim.x must be a multiple of 64 and <= 1024.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[w * NV + k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += lds[i * NV + threadIdx.x];
        lds[16 * NV + threadIdx.x] = s;
    }
    __syncthreads();
}

}  // namespace gamd
