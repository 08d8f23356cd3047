// Covalent terms of the solute on the device (crates/gromos-forces/src/bonded):
// quartic bonds, cosine-harmonic angles, Chebyshev proper dihedrals and
// harmonic impropers, FP64 in both precisions.
//
// Layout: one record per term, terms of a job kind-major (bonds, angles,
// dihedrals, impropers: the order calculate_bonded_forces_ntf adds them,
// bonded/mod.rs:141-185), each kind in topology order.
//   atoms  int4     global atom indices (unused slots -1)
//   prm    double4  bond (k_quartic, r0), angle (k_cosine, cos theta0),
//                   dihedral (K, cos delta, m), improper (K, q0)
//   kind   int32    0 bond, 1 angle, 2 dihedral, 3 improper
// Pass 1 (one lane per term) writes each term's four atom contributions and its
// energy + virial; pass 2 (one lane per atom) sums an atom's contributions from
// a CSR list in term order, kind by kind, exactly as the reference's per-kind
// buffers are filled and then added (ForceEnergy::add, bonded/mod.rs:54-62), so
// the covalent force of an atom carries the reference's rounding.  No atomics.
#pragma once
#include "device_common.hpp"

namespace gamd {

enum BondKind { BD_BOND = 0, BD_ANGLE = 1, BD_DIHEDRAL = 2, BD_IMPROPER = 3 };
constexpr int BD_EV = 10;   // per term: energy, virial[9] (row a, column b)

struct BondedDev {
    const int4* atoms;
    const double4* prm;
    const int32_t* kind;
    double* tf;              // 12 per term: contributions to its atoms 0..3
    double* tev;             // BD_EV per term
    const int32_t* aoff;     // natoms + 1: entries of each gathered atom
    const int32_t* aref;     // entry: term << 2 | slot, in term order
    const int32_t* amap;     // gathered atom -> global atom index (nullptr: identity)
    double* fb;              // 3 per gathered atom: covalent force
    int32_t nterms;
    int32_t natoms;          // gathered atoms
};

namespace bd {

__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 scale(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
// glam DVec3::cross
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ double clamp1(double c) { return c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c); }
// f64::powi with a constant exponent: square-and-multiply from the low bit
__device__ __forceinline__ double powi(double x, int n) {
    double r = 1.0;
    for (;;) {
        if (n & 1) r *= x;
        n >>= 1;
        if (!n) break;
        x *= x;
    }
    return r;
}

struct TermOut {
    d3 f[4];
    double e;
    double vir[9];
};

__device__ __forceinline__ void zero(TermOut& o) {
    for (int s = 0; s < 4; ++s) o.f[s] = {0.0, 0.0, 0.0};
    o.e = 0.0;
    for (int c = 0; c < 9; ++c) o.vir[c] = 0.0;
}

// calculate_bond_forces_quartic (bonds.rs:12-50)
__device__ __forceinline__ void bond(d3 xi, d3 xj, double4 p, TermOut& o) {
    const d3 r = sub(xi, xj);
    const double r2 = dot(r, r);
    const double r0_2 = p.y * p.y;
    const double dr2 = r2 - r0_2;
    o.e = 0.25 * p.x * dr2 * dr2;
    const d3 fv = scale(r, -p.x * dr2);
    o.f[0] = fv;
    o.f[1] = {-fv.x, -fv.y, -fv.z};
    const double rv[3] = {r.x, r.y, r.z}, fa[3] = {fv.x, fv.y, fv.z};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) o.vir[3 * a + c] = rv[a] * fa[c];
}

// calculate_angle_forces, cosine-harmonic (angles.rs:12-75); p.y = cos(theta0)
__device__ __forceinline__ void angle(d3 xi, d3 xj, d3 xk, double4 p, TermOut& o) {
    const d3 rij = sub(xi, xj), rkj = sub(xk, xj);
    const double dij = sqrt(dot(rij, rij)), dkj = sqrt(dot(rkj, rkj));
    if (dij < 1e-10 || dkj < 1e-10) return;
    const double cost = clamp1(dot(rij, rkj) / (dij * dkj));
    const double d_cos = cost - p.y;
    o.e = 0.5 * p.x * d_cos * d_cos;
    const double df = -p.x * d_cos;
    const double a1 = 1.0 / dkj, a2 = cost / dij, a3 = df / dij;
    const double b1 = 1.0 / dij, b2 = cost / dkj, b3 = df / dkj;
    const d3 fi = {(rkj.x * a1 - rij.x * a2) * a3, (rkj.y * a1 - rij.y * a2) * a3, (rkj.z * a1 - rij.z * a2) * a3};
    const d3 fk = {(rij.x * b1 - rkj.x * b2) * b3, (rij.y * b1 - rkj.y * b2) * b3, (rij.z * b1 - rkj.z * b2) * b3};
    o.f[0] = fi;
    o.f[1] = {-(fi.x + fk.x), -(fi.y + fk.y), -(fi.z + fk.z)};
    o.f[2] = fk;
    const double ri[3] = {rij.x, rij.y, rij.z}, rk[3] = {rkj.x, rkj.y, rkj.z};
    const double gi[3] = {fi.x, fi.y, fi.z}, gk[3] = {fk.x, fk.y, fk.z};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) o.vir[3 * a + c] = ri[a] * gi[c] + rk[a] * gk[c];
}

__device__ __forceinline__ void four_atom_virial(d3 rij, d3 rkj, d3 rlj, d3 fi, d3 fk, d3 fl, TermOut& o) {
    const double ri[3] = {rij.x, rij.y, rij.z}, rk[3] = {rkj.x, rkj.y, rkj.z}, rl[3] = {rlj.x, rlj.y, rlj.z};
    const double gi[3] = {fi.x, fi.y, fi.z}, gk[3] = {fk.x, fk.y, fk.z}, gl[3] = {fl.x, fl.y, fl.z};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o.vir[3 * a + b] = ri[a] * gi[b] + rk[a] * gk[b] + rl[a] * gl[b];
}

// calculate_dihedral_forces (dihedrals.rs:19-130); p = (K, cos delta, m)
__device__ __forceinline__ void dihedral(d3 xi, d3 xj, d3 xk, d3 xl, double4 p, TermOut& o) {
    const d3 rij = sub(xi, xj), rkj = sub(xk, xj), rkl = sub(xk, xl);
    const double dkj2 = dot(rkj, rkj);
    if (dkj2 < 1e-10) return;
    const double frim = dot(rij, rkj) / dkj2, frln = dot(rkl, rkj) / dkj2;
    const d3 rim = sub(rij, scale(rkj, frim));
    const d3 rln = sub(scale(rkj, frln), rkl);
    const double dim = sqrt(dot(rim, rim)), dln = sqrt(dot(rln, rln));
    if (dim < 1e-10 || dln < 1e-10) return;
    const double c = clamp1(dot(rim, rln) / (dim * dln));
    double cm, dcm;
    switch ((int)p.z) {
        case 0: cm = 0.0; dcm = 0.0; break;
        case 1: cm = c; dcm = 1.0; break;
        case 2: cm = 2.0 * c * c - 1.0; dcm = 4.0 * c; break;
        case 3: cm = 4.0 * powi(c, 3) - 3.0 * c; dcm = 12.0 * c * c - 3.0; break;
        case 4: cm = 8.0 * powi(c, 4) - 8.0 * c * c + 1.0; dcm = 32.0 * powi(c, 3) - 16.0 * c; break;
        case 5:
            cm = 16.0 * powi(c, 5) - 20.0 * powi(c, 3) + 5.0 * c;
            dcm = 80.0 * powi(c, 4) - 60.0 * c * c + 5.0;
            break;
        default:   // 6 (the layout admits 0..6 only)
            cm = 32.0 * powi(c, 6) - 48.0 * powi(c, 4) + 18.0 * c * c - 1.0;
            dcm = 192.0 * powi(c, 5) - 192.0 * powi(c, 3) + 36.0 * c;
            break;
    }
    const double K = p.x, cospd = p.y;
    o.e = K * (1.0 + cospd * cm);
    const double ki = -K * cospd * dcm / dim, kl = -K * cospd * dcm / dln;
    const double kj1 = frim - 1.0, kj2 = frln;
    const double i1 = ki / dln, i2 = ki * c / dim, l1 = kl / dim, l2 = kl * c / dln;
    const d3 fi = {rln.x * i1 - rim.x * i2, rln.y * i1 - rim.y * i2, rln.z * i1 - rim.z * i2};
    const d3 fl = {rim.x * l1 - rln.x * l2, rim.y * l1 - rln.y * l2, rim.z * l1 - rln.z * l2};
    const d3 fj = {fi.x * kj1 - fl.x * kj2, fi.y * kj1 - fl.y * kj2, fi.z * kj1 - fl.z * kj2};
    const d3 fk = {-(fi.x + fj.x + fl.x), -(fi.y + fj.y + fl.y), -(fi.z + fj.z + fl.z)};
    o.f[0] = fi;
    o.f[1] = fj;
    o.f[2] = fk;
    o.f[3] = fl;
    four_atom_virial(rij, rkj, sub(xl, xj), fi, fk, fl, o);
}

// calculate_improper_dihedral_forces (improper.rs:12-105); p = (K, q0)
__device__ __forceinline__ void improper(d3 xi, d3 xj, d3 xk, d3 xl, double4 p, TermOut& o) {
    const double PI = 3.14159265358979323846;
    const d3 rkj = sub(xk, xj), rij = sub(xi, xj), rkl = sub(xk, xl);
    const d3 rmj = cross(rij, rkj), rnk = cross(rkj, rkl);
    const double dkj2 = dot(rkj, rkj), dmj2 = dot(rmj, rmj), dnk2 = dot(rnk, rnk);
    const double dkj = sqrt(dkj2), dmj = sqrt(dmj2), dnk = sqrt(dnk2);
    if (dmj < 1e-10 || dnk < 1e-10) return;
    const double acs = clamp1(dot(rmj, rnk) / (dmj * dnk));
    double zeta = acos(acs);
    if (dot(rij, rnk) < 0.0) zeta = -zeta;
    const double K = p.x, q0 = p.y;
    // the reference's unbounded while loops; zeta is within [-pi, pi] and a
    // finite q0 needs few turns, a non-finite one none
    for (int t = 0; t < 64 && zeta < q0 - PI; ++t) zeta += 2.0 * PI;
    for (int t = 0; t < 64 && zeta > q0 + PI; ++t) zeta -= 2.0 * PI;
    const double dz = zeta - q0;
    o.e = 0.5 * K * dz * dz;
    double ki = -K * dz * dkj, kl = -ki;
    ki = dmj2 < 1.0e-10 * dkj2 ? 0.0 : ki / dmj2;
    kl = dnk2 < 1.0e-10 * dkj2 ? 0.0 : kl / dnk2;
    const double kj1 = dot(rij, rkj) / dkj2 - 1.0, kj2 = dot(rkl, rkj) / dkj2;
    const d3 fi = scale(rmj, ki), fl = scale(rnk, kl);
    const d3 fj = {fi.x * kj1 - fl.x * kj2, fi.y * kj1 - fl.y * kj2, fi.z * kj1 - fl.z * kj2};
    const d3 fk = {-(fi.x + fj.x + fl.x), -(fi.y + fj.y + fl.y), -(fi.z + fj.z + fl.z)};
    o.f[0] = fi;
    o.f[1] = fj;
    o.f[2] = fk;
    o.f[3] = fl;
    four_atom_virial(rij, rkj, sub(xl, xj), fi, fk, fl, o);
}

}  // namespace bd

// pass 1 over x (3 doubles per global atom), pass 2 into bd.fb; then the
// per-term energies / virials stay in bd.tev for a reduction
void launch_bonded(const BondedDev& bd, const double* x, hipStream_t s);
// f[3 amap[w] + c] = fb[3 w + c] + f[3 amap[w] + c] (covalent + nonbonded,
// forcefield.rs:966)
void launch_bonded_add(const BondedDev& bd, double* f, hipStream_t s);

}  // namespace gamd
