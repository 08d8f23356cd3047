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

// Boundary: kind 0 = vacuum, 1 = rectangular.  invL = 1.0 / L computed once on
// the host (DVec3::recip, math.rs:49-55).
struct Box {
    double L[3];
    double invL[3];
    int32_t kind;
    int32_t pad;
};

__device__ __forceinline__ double nim1(double d, double L, double invL, int kind) {
    return kind ? d - L * round(d * invL) : d;
}

__device__ __forceinline__ d3 nim(const Box& b, d3 ri, d3 rj) {
    d3 r{ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    if (b.kind) {
        r.x = r.x - b.L[0] * round(r.x * b.invL[0]);
        r.y = r.y - b.L[1] * round(r.y * b.invL[1]);
        r.z = r.z - b.L[2] * round(r.z * b.invL[2]);
    }
    return r;
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
