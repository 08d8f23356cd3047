// Host driver of the device-resident systems: job images, batch layout, the
// pair-list rebuild cadence, Forcefield::apply and the md step sequence.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "host.hpp"

using namespace gamd;

namespace gamd {

// box: 3 edge lengths (rectangular), or the vectors a, b, c (9 values, triclinic)
Box make_box(int32_t boundary, const double* box) {
    Box b{};
    b.kind = boundary == 1 || boundary == 2 ? boundary : 0;
    if (b.kind == 2) {
        for (int a = 0; a < 3; ++a) {
            b.a[a] = box[a];
            b.b[a] = box[3 + a];
            b.c[a] = box[6 + a];
        }
        b.L[0] = b.a[0];
        b.L[1] = b.b[1];
        b.L[2] = b.c[2];
        for (int a = 0; a < 3; ++a) b.invL[a] = 1.0 / b.L[a];
        b.tr[0] = b.c[1] / b.c[2];
        b.tr[1] = b.b[0] / b.b[1];
        b.tr[2] = b.c[0] / b.c[2];
        return b;
    }
    for (int a = 0; a < 3; ++a) {
        b.L[a] = b.kind ? box[a] : 0.0;
        b.invL[a] = b.kind ? 1.0 / box[a] : 0.0;   // DVec3::recip
    }
    return b;
}

int check_box(int32_t boundary, const double* box) {
    if (boundary < 0 || boundary > 2)
        return fail(GROMOS_AMD_EINVAL, "boundary must be 0 (vacuum), 1 (rectangular) or 2 (triclinic)");
    if (boundary == 1) {
        if (!box) return fail(GROMOS_AMD_EINVAL, "rectangular boundary needs a box");
        for (int a = 0; a < 3; ++a)
            if (!(box[a] > 0.0)) return fail(GROMOS_AMD_EINVAL, "box edge %d must be > 0", a);
    }
    if (boundary == 2) {
        // the GROMOS lower-triangular convention a = (ax, 0, 0), b = (bx, by, 0),
        // c = (cx, cy, cz) that Triclinic::nearest_image requires (math.rs:97-100)
        if (!box) return fail(GROMOS_AMD_EINVAL, "triclinic boundary needs the box vectors a, b, c");
        if (box[1] != 0.0 || box[2] != 0.0 || box[5] != 0.0)
            return fail(GROMOS_AMD_EINVAL, "triclinic box vectors must be lower triangular (a = (ax,0,0), b = (bx,by,0))");
        if (!(box[0] > 0.0) || !(box[4] > 0.0) || !(box[8] > 0.0))
            return fail(GROMOS_AMD_EINVAL, "triclinic box diagonal ax, by, cz must be > 0");
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(box[k])) return fail(GROMOS_AMD_EINVAL, "triclinic box vectors must be finite");
    }
    return 0;
}

// the box the C ABI's params carry for each boundary kind
static const double* params_box(const gromos_amd_params& P) { return P.boundary == 2 ? P.box_matrix : P.box; }

// cells per axis: the largest count whose edge still exceeds half the list
// cutoff (with a relative margin so that rounding in the wrapped coordinates can
// never push an in-range pair three cells apart; the search takes two cells on
// each side); 1 when the box is smaller.
static int32_t cells_for(double L, double rl) {
    if (!(rl > 0.0)) return 1;
    int64_t n = (int64_t)std::floor(L / (0.5 * rl * (1.0 + 1e-9) + 1e-12));
    if (n < 1) n = 1;
    if (n > 1024) n = 1024;
    return (int32_t)n;
}

void rebox_image(JobImage& im, const Box& box) {
    JobDev& J = im.J;
    J.box = box;
    if (box.kind == 2) return;   // one cell, CG-grouped loops off, whatever the size
    double hm = 1e300;
    for (int a = 0; a < 3; ++a) {
        J.ncell[a] = cells_for(box.L[a] * 0.95, 2.0 * J.cell_min);
        hm = std::min(hm, 0.5 * box.L[a]);
    }
    J.use_cg_grouped = hm > J.cgg_limit;
}

// device bytes per padded atom / per charge group held by a batch (the arrays
// of gromos_amd_system), used for admission by footprint
constexpr int64_t BYTES_PER_ATOM = 2 * 2 * 24 + 7 * 24 + 32 + 16 + 8 + 4 + 8 + 8 + 4 + 12 + 16 + 41 +
                                   NSLOT * ROW * 8 * 4 / BLOCK + 16;
constexpr int64_t BYTES_PER_CG = 2 * 24 + 16 + 16 + 16 + 4 * 8 + 2 * 8 + 2 * 4 + 8;

// setup_lincs (constraints.rs:1168-1220): sd[k] = 1 / sqrt(1/m_i + 1/m_j); two
// constraints a < b sharing an atom (checked first-first, second-first,
// first-second, second-second) are coupled with (1/m_common) sd[a] sd[b],
// negated when the atom is the first of both or the second of both; each
// constraint's coupled ones ascend.  ij: atom pairs; inv_m(atom).
template <class F>
void lincs_setup(int32_t nc, const int32_t* ij, F inv_m, std::vector<double>& sd,
                 std::vector<std::vector<std::pair<int32_t, double>>>& cpl) {
    sd.assign((size_t)nc, 0.0);
    cpl.assign((size_t)nc, {});
    std::unordered_map<int32_t, std::vector<int32_t>> by_atom;
    for (int32_t k = 0; k < nc; ++k) {
        sd[k] = 1.0 / std::sqrt(inv_m(ij[2 * k]) + inv_m(ij[2 * k + 1]));
        by_atom[ij[2 * k]].push_back(k);
        if (ij[2 * k + 1] != ij[2 * k]) by_atom[ij[2 * k + 1]].push_back(k);
    }
    for (int32_t k = 0; k < nc; ++k) {
        std::vector<int32_t> other;
        for (int s = 0; s < 2; ++s)
            for (int32_t c : by_atom[ij[2 * k + s]])
                if (c != k) other.push_back(c);
        std::sort(other.begin(), other.end());
        other.erase(std::unique(other.begin(), other.end()), other.end());
        for (int32_t c : other) {
            const int32_t a = std::min(k, c), b = std::max(k, c);
            const int32_t ai = ij[2 * a], aj = ij[2 * a + 1], bi = ij[2 * b], bj = ij[2 * b + 1];
            const int32_t common = ai == bi ? ai : (aj == bi ? aj : (ai == bj ? ai : aj));
            double v = inv_m(common) * sd[a] * sd[b];
            if (ai == bi || aj == bj) v *= -1.0;
            cpl[k].push_back({c, v});
        }
    }
}

int make_image(const gromos_amd_topology& T, const gromos_amd_params& P, int j, JobImage& im) {
    if (int rc = check_box(P.boundary, params_box(P))) return rc;
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
    if (P.rcutl < P.rcutp || !(P.rcutl > 0.0)) return fail(GROMOS_AMD_EINVAL, "job %d: need 0 < rcutp <= rcutl", j);
    if (!(P.dt > 0.0)) return fail(GROMOS_AMD_EINVAL, "job %d: dt must be > 0", j);
    im = JobImage{};
    im.P = P;
    if (P.solvent_shake && P.solvent_lincs > 0)
        return fail(GROMOS_AMD_EINVAL, "job %d: the solvent is constrained by SHAKE or by LINCS, not both", j);
    if (P.solute_lincs < 0 || P.solvent_lincs < 0 || P.solute_lincs > 64 || P.solvent_lincs > 64)
        return fail(GROMOS_AMD_EINVAL, "job %d: LINCS orders must be 0..64", j);
    if (P.solvent_shake || P.solvent_lincs > 0) im.P.use_settle = 0;   // SHAKE or LINCS instead
    im.n = n;
    im.n_cg = T.n_cg;
    im.n_types = T.n_types;
    im.solvent = nsolv_atoms > 0;
    JobDev& J = im.J;
    J.box = make_box(P.boundary, params_box(P));
    gromos_amd_crf c;
    gromos_amd_crf_new(P.rcutl, 1.0, P.epsrf, P.kappa, &c);   // md.rs:326-329, 815
    J.crf = to_crf(&c);
    J.fpepsi = P.four_pi_eps_i;
    double s = P.rcutp + P.skin, l = P.rcutl + P.skin;
    J.c2s = s * s;   // (rc + skin).powi(2), pairlist.rs:192-193
    J.c2l = l * l;
    J.n_types = T.n_types;
    J.aps = aps;
    J.n_atoms = n;
    J.n_solute_atoms = T.n_solute_atoms;
    J.n_cg = T.n_cg;
    J.n_solute_cg = T.n_solute_cg;
    J.twin = P.rcutp < P.rcutl - 1e-10;
    J.dt = P.dt;
    J.dist_oh = T.dist_oh;
    J.dist_hh = T.dist_hh;
    J.stop = INT32_MAX;
    static_assert(NHC_MAX == GROMOS_AMD_NHC_MAX, "chain slots per job");
    if (P.thermostat < 0 || P.thermostat > 2) return fail(GROMOS_AMD_EINVAL, "job %d: thermostat must be 0, 1 or 2", j);
    if (P.thermostat == 2 && (P.nhc < 1 || P.nhc > GROMOS_AMD_NHC_MAX))
        return fail(GROMOS_AMD_EINVAL, "job %d: Nose-Hoover chain length %d outside 1..%d", j, P.nhc,
                    GROMOS_AMD_NHC_MAX);
    J.thermo = P.thermostat;
    J.nhc = P.thermostat == 2 ? P.nhc : 0;
    if (P.nticom < 0 || P.nticom > 2) return fail(GROMOS_AMD_EINVAL, "job %d: nticom must be 0, 1 or 2", j);
    J.nticom = P.nticom;
    J.nscm = P.nscm;
    J.t_ref = P.t_ref;
    J.tau_t = P.tau_t;
    J.dof_t = P.dof_t;
    if (P.barostat < 0 || P.barostat > 1) return fail(GROMOS_AMD_EINVAL, "job %d: barostat must be 0 or 1", j);
    if (P.barostat) {
        if (P.boundary != 1 && P.boundary != 2)
            return fail(GROMOS_AMD_EINVAL, "job %d: the barostat needs a periodic box", j);
        if (P.virial < 0 || P.virial > 2 || !(P.tau_p > 0.0))
            return fail(GROMOS_AMD_EINVAL, "job %d: barostat virial must be 0, 1 or 2 and tau_p > 0", j);
    }
    J.baro = P.barostat;
    J.vtype = P.barostat ? P.virial : 0;
    J.p_ref = P.p_ref;
    J.comp = P.compressibility;
    J.tau_p = P.tau_p;
    J.cgg_limit = P.rcutl + P.skin + 0.3;
    J.cell_min = 0.5 * (P.rcutl + P.skin);
    if (J.baro && J.vtype == 2) {
        // pressure groups: the solute's, then each solvent molecule (core topology.rs:1242-1263)
        im.pg_start.assign(1, 0);
        int32_t prev = -1;
        for (int32_t g = 0; g < T.n_pressure_groups; ++g) {
            const int32_t last = T.pressure_group_last ? T.pressure_group_last[g] : -1;
            if (last <= prev || last >= T.n_solute_atoms)
                return fail(GROMOS_AMD_EINVAL, "job %d: pressure group %d must end after %d and before %d", j, g,
                            prev, T.n_solute_atoms);
            im.pg_start.push_back(last + 1);
            prev = last;
        }
        if (T.n_pressure_groups > 0 && prev != T.n_solute_atoms - 1)
            return fail(GROMOS_AMD_EINVAL, "job %d: the pressure groups must cover the solute", j);
        if (T.n_pressure_groups == 0 && T.n_solute_atoms > 0)
            return fail(GROMOS_AMD_EINVAL, "job %d: a molecular virial needs the solute's pressure groups", j);
        for (int32_t a = T.n_solute_atoms + T.atoms_per_solvent; a <= n && T.atoms_per_solvent > 0;
             a += T.atoms_per_solvent)
            im.pg_start.push_back(a);
    }
    // use_cg_grouped_solute (forcefield.rs:271-283; triclinic: off)
    if (J.box.kind == 0) J.use_cg_grouped = 1;
    else if (J.box.kind == 2) J.use_cg_grouped = 0;
    else {
        double hm = std::min(std::min(0.5 * P.box[0], 0.5 * P.box[1]), 0.5 * P.box[2]);
        J.use_cg_grouped = hm > P.rcutl + P.skin + 0.3;
    }
    // vacuum and triclinic jobs: one cell, every group a candidate of every other
    // (the reference's O(N^2) standard search for both, pairlist.rs:521-532)
    // a barostat's box: cells with 5 % to spare, so that the grid holds while the
    // box shrinks that much (the step checks it and stops the job beyond)
    for (int a = 0; a < 3; ++a)
        J.ncell[a] = J.box.kind == 1 ? cells_for(P.box[a] * (J.baro ? 0.95 : 1.0), l) : 1;
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
                im.has_pert = true;
                if (a >= T.n_solute_atoms) return fail(GROMOS_AMD_EINVAL, "job %d: only solute atoms can be perturbed", j);
                if (!T.pert_a_iac || !T.pert_b_iac || !T.pert_a_q || !T.pert_b_q || !T.pert_alpha_lj || !T.pert_alpha_crf)
                    return fail(GROMOS_AMD_EINVAL, "job %d: incomplete perturbation arrays", j);
                if (T.pert_a_iac[a] < 0 || T.pert_a_iac[a] >= T.n_types || T.pert_b_iac[a] < 0 || T.pert_b_iac[a] >= T.n_types)
                    return fail(GROMOS_AMD_EINVAL, "job %d: perturbed type of atom %d out of range", j, a);
            }
    }
    J.has_pert = im.has_pert ? 1 : 0;
    J.settle = (im.P.use_settle && nsolv_atoms > 0 && aps == 3) ? 1 : 0;
    // SHAKE: the solvent template and the solute's constraints
    J.sh_tol = P.shake_tolerance > 0.0 ? P.shake_tolerance : 1e-4;
    J.shake_sv = 0;
    J.n_sc = 0;
    J.lincs_u = P.solute_lincs;
    J.lincs_sv = 0;
    if ((P.solvent_shake || P.solvent_lincs > 0) && nsolv_atoms > 0) {
        const int32_t nc = T.n_solvent_constraints;
        if (nc < 1 || nc > 6 || !T.solvent_constraint_ij || !T.solvent_constraint_len)
            return fail(GROMOS_AMD_EINVAL, "job %d: solvent SHAKE / LINCS needs 1..6 template constraints", j);
        for (int32_t c = 0; c < nc; ++c) {
            const int32_t a = T.solvent_constraint_ij[2 * c], b = T.solvent_constraint_ij[2 * c + 1];
            if (a < 0 || b < 0 || a >= aps || b >= aps || a == b || !(T.solvent_constraint_len[c] > 0.0))
                return fail(GROMOS_AMD_EINVAL, "job %d: solvent constraint %d out of range", j, c);
            J.sc_ij[2 * c] = a;
            J.sc_ij[2 * c + 1] = b;
            J.sc_len[c] = T.solvent_constraint_len[c];
        }
        J.n_sc = nc;
        if (P.solvent_lincs > 0) {
            // the template's LINCS matrix from the first molecule's masses (constraints.rs:1130-1137)
            double inv_m[4] = {0.0, 0.0, 0.0, 0.0};
            for (int32_t a = 0; a < aps; ++a) {
                const double m = T.mass[T.n_solute_atoms + a];
                inv_m[a] = m > 0.0 ? 1.0 / m : 0.0;
            }
            std::vector<double> sd;
            std::vector<std::vector<std::pair<int32_t, double>>> cpl;
            lincs_setup(nc, J.sc_ij, [&](int32_t a) { return inv_m[a]; }, sd, cpl);
            for (int32_t k = 0; k < 6; ++k) {
                J.lc_sdiag[k] = k < nc ? sd[k] : 0.0;
                J.lc_cpl[k] = 0;
                for (int32_t c = 0; c < 6; ++c) J.lc_coef[6 * k + c] = 0.0;
            }
            for (int32_t k = 0; k < nc; ++k)
                for (const auto& e : cpl[k]) {
                    J.lc_cpl[k] |= 1 << e.first;
                    J.lc_coef[6 * k + e.first] = e.second;
                }
            J.lincs_sv = P.solvent_lincs;
        } else {
            J.shake_sv = 1;
        }
    }
    im.shake.clear();
    for (int32_t c = 0; c < T.n_solute_constraints; ++c) {
        const int32_t a = T.solute_constraint_ij ? T.solute_constraint_ij[2 * c] : -1;
        const int32_t b = T.solute_constraint_ij ? T.solute_constraint_ij[2 * c + 1] : -1;
        const double l = T.solute_constraint_len ? T.solute_constraint_len[c] : 0.0;
        if (a < 0 || b < 0 || a >= T.n_solute_atoms || b >= T.n_solute_atoms || a == b || !(l > 0.0))
            return fail(GROMOS_AMD_EINVAL, "job %d: solute constraint %d out of range", j, c);
        im.shake.push_back({a, b, l});
    }
    // from_lambda(l, l, l, l, 1, 1, 1, 1, NLAM) (forcefield.rs:767-777)
    {
        gromos_amd_lambda L;
        gromos_amd_lambda_from(P.lambda, P.lambda, P.lambda, P.lambda, 1.0, 1.0, 1.0, 1.0, P.nlam, &L);
        std::memcpy(&J.lp, &L, sizeof(L));
    }
    if (im.P.use_settle && nsolv_atoms > 0) {
        if (aps != 3) return fail(GROMOS_AMD_EINVAL, "SETTLE needs 3-atom solvent (settle_algorithm.rs:43-45)");
        if (T.mass[T.n_solute_atoms + 1] != T.mass[T.n_solute_atoms + 2])
            return fail(GROMOS_AMD_EINVAL, "SETTLE needs equal H masses (settle_algorithm.rs:46-51)");
        for (int32_t a = T.n_solute_atoms; a + 3 <= n; a += 3) im.mol_first.push_back(a);
    }
    im.cg_start.assign(T.cg_start, T.cg_start + T.n_cg + 1);
    im.q.assign(T.charge, T.charge + n);
    im.mass.assign(T.mass, T.mass + n);
    im.iac.assign(T.iac, T.iac + n);
    im.lj.assign(T.lj, T.lj + (int64_t)T.n_types * T.n_types * 4);
    im.excl_off.reserve(n + 1);
    im.o14_off.reserve(n + 1);
    im.o14s_off.reserve(n + 1);
    im.excl_off.push_back(0);
    im.o14_off.push_back(0);
    im.o14s_off.push_back(0);
    std::vector<std::vector<int32_t>> sym(n);
    for (int a = 0; a < n; ++a) {
        for (int k = T.excl_off[a]; k < T.excl_off[a + 1]; ++k) {
            int32_t b = T.excl[k];
            if (b < 0 || b >= n) return fail(GROMOS_AMD_EINVAL, "job %d: exclusion index out of range", j);
            im.excl.push_back(b);
        }
        for (int k = T.o14_off[a]; k < T.o14_off[a + 1]; ++k) {
            int32_t b = T.o14[k];
            if (b < 0 || b >= n) return fail(GROMOS_AMD_EINVAL, "job %d: 1-4 index out of range", j);
            im.o14.push_back(b);
            if (b > a) {  // excluded.rs:148-151 evaluates j > i only
                sym[a].push_back(b | (int32_t)0x80000000u);
                sym[b].push_back(a);
            }
        }
        im.excl_off.push_back((int32_t)im.excl.size());
        im.o14_off.push_back((int32_t)im.o14.size());
    }
    for (int a = 0; a < n; ++a) {
        for (int32_t e : sym[a]) im.o14s.push_back(e);
        im.o14s_off.push_back((int32_t)im.o14s.size());
    }
    if (im.has_pert) {
        im.pflag.assign(T.pert_flag, T.pert_flag + n);
        im.pa_iac.assign(T.pert_a_iac, T.pert_a_iac + n);
        im.pb_iac.assign(T.pert_b_iac, T.pert_b_iac + n);
        im.pa_q.assign(T.pert_a_q, T.pert_a_q + n);
        im.pb_q.assign(T.pert_b_q, T.pert_b_q + n);
        im.palj.assign(T.pert_alpha_lj, T.pert_alpha_lj + n);
        im.pacrf.assign(T.pert_alpha_crf, T.pert_alpha_crf + n);
        for (int k = 0; k < T.n_pert_atom_pairs; ++k) {
            const int32_t* e = T.pert_atom_pairs + 4 * k;
            if (e[0] < 0 || e[0] >= n || e[1] < 0 || e[1] >= n) return fail(GROMOS_AMD_EINVAL, "job %d: atom pair out of range", j);
            // like the perturbed atoms themselves, PERTATOMPAIR entries are solute
            // atoms (the solvent takes no excluded / 1-4 / atom-pair forces here)
            if (e[0] >= T.n_solute_atoms || e[1] >= T.n_solute_atoms)
                return fail(GROMOS_AMD_EINVAL, "job %d: PERTATOMPAIR atoms must be solute atoms", j);
            for (int c = 0; c < 4; ++c) im.papairs.push_back(e[c]);
        }
    }
    if (T.bonded) {
        if (int rc = bonded_terms(*T.bonded, n, j, im.bonded)) return rc;
        // the gathered atoms of a batch are its solute atoms
        for (const int4& a : im.bonded.atoms)
            for (int32_t v : {a.x, a.y, a.z, a.w})
                if (v >= T.n_solute_atoms)
                    return fail(GROMOS_AMD_EINVAL, "job %d: covalent terms must join solute atoms", j);
        if (im.bonded.atoms.size() >= (size_t)(INT32_MAX >> 3))
            return fail(GROMOS_AMD_EINVAL, "job %d: too many covalent terms", j);
    }
    J.bd_off = 0;
    for (int k = 0; k < 4; ++k) J.bd_n[k] = im.bonded.n[k];
    // footprint: per-atom and per-group arrays, the full neighbour lists
    // (partners within the list cutoff at this job's group density, both
    // directions, 4 bytes each, with the buffers' growth policy: 25 % slack,
    // doubled when outgrown) and the energy records of md calls of up to 1,000
    // steps (doubled likewise)
    int64_t npad = align_up(n, BLOCK);
    double ncg = T.n_cg;
    double partners = ncg;
    if (J.box.kind) {
        double vol = J.box.L[0] * J.box.L[1] * J.box.L[2];   // a lower-triangular box: the diagonal's product
        double rr = l + 0.3;
        partners = std::min(ncg, 4.18879 * rr * rr * rr * ncg / vol);
    }
    int64_t cells = (int64_t)J.ncell[0] * J.ncell[1] * J.ncell[2];
    im.list_entries = (int64_t)(ncg * partners * 1.25);
    im.footprint = npad * BYTES_PER_ATOM + (int64_t)T.n_cg * BYTES_PER_CG + 8 * cells +
                   (int64_t)(ncg * partners * 4.0 * 2.5) + (int64_t)(im.excl.size() + im.o14.size() * 2) * 4 +
                   (int64_t)(2.5 * 1000 * REC * sizeof(double)) + 4096 +
                   (int64_t)im.bonded.atoms.size() * (16 + 32 + 4 + 8 * (12 + BD_EV) + 4 * 4) * 5 / 4;
    return 0;
}

// ---------------------------------------------------------------------------
// batch layout
// ---------------------------------------------------------------------------
int layout_batch(gromos_amd_system* S, const std::vector<const JobImage*>& imgs, int32_t precision) {
    if (imgs.empty()) return fail(GROMOS_AMD_EINVAL, "layout: no jobs");
    if (!S->stream) HIPCHK(hipStreamCreate(&S->stream));
    if (!S->side) {
        HIPCHK(hipStreamCreate(&S->side));
        HIPCHK(hipEventCreateWithFlags(&S->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&S->ev_join, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&S->ctl_ev, hipEventDisableTiming));
    }
    HIPCHK(hipStreamSynchronize(S->stream));   // host vectors below feed async copies
    int njobs = (int)imgs.size();
    const JobImage& I0 = *imgs[0];
    int32_t aps0 = 0;
    for (const JobImage* im : imgs)
        if (im->solvent) {
            if (aps0 && im->J.aps != aps0) return fail(GROMOS_AMD_EINVAL, "jobs of one batch share the solvent molecule size");
            aps0 = im->J.aps;
        }
    for (const JobImage* im : imgs)
        if (im->P.nsnb != I0.P.nsnb) return fail(GROMOS_AMD_EINVAL, "jobs of one batch share the rebuild cadence");
    S->precision = precision;
    S->njobs = njobs;
    S->host_job_pairs.assign(JOB_CTL * (size_t)njobs, 0);
    S->jobs.clear();
    S->job_n_atoms.clear();
    const auto t_concat = std::chrono::steady_clock::now();
    // pass 1 (serial): each job's offsets in every concatenated array
    struct Slice {
        int64_t atom, cg, lj, excl, o14, o14s, molf, pap, sol, bd;
    };
    std::vector<Slice> sl(njobs + 1);
    int64_t cell_off = 0;
    bool any_pert = false, twin = false, any_thermo = false, any_com = false, any_baro = false;
    bool any_solute_shake = false, any_solvent_pass = false, any_lincs_u = false;
    S->host_pg.clear();
    S->host_sh_ij.clear();
    S->host_sh_len.clear();
    S->host_lc_sdiag.clear();
    S->host_lc_cpl.clear();
    S->host_lc_cidx.clear();
    S->host_lc_coef.clear();
    int32_t use_settle = 0, do_virial = 0;
    for (int j = 0; j < njobs; ++j) {
        const JobImage& im = *imgs[j];
        const Slice& o = sl[j];
        Slice& n = sl[j + 1];
        n.atom = o.atom + align_up(im.n, BLOCK);
        n.cg = o.cg + im.n_cg;
        n.lj = o.lj + (int64_t)im.lj.size();
        n.excl = o.excl + (int64_t)im.excl.size();
        n.o14 = o.o14 + (int64_t)im.o14.size();
        n.o14s = o.o14s + (int64_t)im.o14s.size();
        n.molf = o.molf + (im.P.use_settle ? (int64_t)im.mol_first.size() : 0);
        n.pap = o.pap + (int64_t)im.papairs.size();
        n.sol = o.sol + im.J.n_solute_atoms;
        n.bd = o.bd + (int64_t)im.bonded.atoms.size();
        JobDev J = im.J;
        J.bd_off = (int32_t)o.bd;
        J.atom_off = (int32_t)o.atom;
        J.cg_off = (int32_t)o.cg;
        J.lj_off = (int32_t)o.lj;
        J.cell_off = (int32_t)cell_off;
        J.stop = INT32_MAX;
        J.sol_off = (int32_t)o.sol;
        cell_off += (int64_t)J.ncell[0] * J.ncell[1] * J.ncell[2];
        any_pert = any_pert || im.has_pert;
        twin = twin || J.twin;
        any_thermo = any_thermo || J.thermo;
        any_com = any_com || J.nticom >= 1 || J.nscm != 0;
        any_baro = any_baro || J.baro;
        J.pg_off = (int32_t)S->host_pg.size();
        J.n_pg = im.pg_start.empty() ? 0 : (int32_t)im.pg_start.size() - 1;
        for (int32_t x : im.pg_start) S->host_pg.push_back((int32_t)o.atom + x);
        J.sh_off = (int32_t)S->host_sh_len.size();
        J.sh_n = (int32_t)im.shake.size();
        for (const auto& c : im.shake) {
            S->host_sh_ij.push_back((int32_t)o.atom + c.i);
            S->host_sh_ij.push_back((int32_t)o.atom + c.j);
            S->host_sh_len.push_back(c.len);
        }
        if (J.lincs_u > 0 && J.sh_n > 0) {
            // LINCS matrix of the job's solute constraints, rows beside sh_ij / sh_len
            std::vector<int32_t> ij;
            for (const auto& c : im.shake) {
                ij.push_back(c.i);
                ij.push_back(c.j);
            }
            std::vector<double> sd;
            std::vector<std::vector<std::pair<int32_t, double>>> cpl;
            lincs_setup(J.sh_n, ij.data(),
                        [&](int32_t a) { return im.mass[a] > 0.0 ? 1.0 / im.mass[a] : 0.0; }, sd, cpl);
            for (int32_t k = 0; k < J.sh_n; ++k) {
                S->host_lc_sdiag.push_back(sd[k]);
                S->host_lc_cpl.push_back((int32_t)S->host_lc_cidx.size());
                S->host_lc_cpl.push_back((int32_t)cpl[k].size());
                for (const auto& e : cpl[k]) {
                    S->host_lc_cidx.push_back(e.first);
                    S->host_lc_coef.push_back(e.second);
                }
            }
            any_lincs_u = true;
        } else {
            for (int32_t k = 0; k < J.sh_n; ++k) {
                S->host_lc_sdiag.push_back(0.0);
                S->host_lc_cpl.push_back(0);
                S->host_lc_cpl.push_back(0);
            }
            J.lincs_u = 0;
        }
        any_solute_shake = any_solute_shake || J.sh_n > 0;
        any_solvent_pass = any_solvent_pass || J.lincs_sv > 0 || J.shake_sv;
        use_settle |= im.P.use_settle && !im.mol_first.empty();
        do_virial |= im.P.do_virial;
        S->jobs.push_back(J);
        S->job_n_atoms.push_back(im.n);
    }
    const Slice& T = sl[njobs];
    const int64_t npad = T.atom, ncg = T.cg;
    if (npad >= (int64_t)ENT_IDX) return fail(GROMOS_AMD_EINVAL, "too many atoms");
    if (ncg >= (int64_t)ENT_IDX) return fail(GROMOS_AMD_EINVAL, "too many charge groups");
    if (T.excl >= INT32_MAX || T.o14 >= INT32_MAX || T.o14s >= INT32_MAX || cell_off >= INT32_MAX ||
        T.bd >= (INT32_MAX >> 3))
        return fail(GROMOS_AMD_EINVAL, "batch too large");
    // the arrays live in the batch's pinned staging arena (all written below)
    const size_t np1 = (size_t)npad + 1, pn = any_pert ? (size_t)npad : 0;
    const size_t n_pap = (size_t)T.pap / 2;   // two index entries per PERTATOMPAIR entry at most
    const size_t bytes = 4 * (3 * (size_t)ncg + 2 * (size_t)npad + 3 * np1 + T.excl + T.o14 + T.o14s + T.molf +
                              T.pap + T.sol) +
                         (size_t)T.bd * (16 + 32 + 4) +
                         8 * (3 * (size_t)npad + T.lj) + pn * (1 + 4 * 2 + 8 * 4) +
                         (any_pert ? (size_t)ncg + 4 * (np1 + n_pap) : 0);
    if (int rc = S->host.reset(bytes, 40)) return rc;
    auto cg_start = S->host.take<int32_t>(ncg), cg_n = S->host.take<int32_t>(ncg), cg_job = S->host.take<int32_t>(ncg);
    auto atom_job = S->host.take<int32_t>(npad), iac = S->host.take<int32_t>(npad);
    auto excl_off = S->host.take<int32_t>(np1), excl = S->host.take<int32_t>(T.excl);
    auto o14_off = S->host.take<int32_t>(np1), o14 = S->host.take<int32_t>(T.o14);
    auto o14s_off = S->host.take<int32_t>(np1), o14s = S->host.take<int32_t>(T.o14s);
    auto mol_first = S->host.take<int32_t>(T.molf), papairs = S->host.take<int32_t>(T.pap);
    auto sol_atom = S->host.take<int32_t>(T.sol);
    auto bd_atoms = S->host.take<int4>(T.bd);
    auto bd_prm = S->host.take<double4>(T.bd);
    auto bd_kind = S->host.take<int32_t>(T.bd);
    auto q = S->host.take<double>(npad), mass = S->host.take<double>(npad), inv_mass = S->host.take<double>(npad);
    auto lj = S->host.take<double>(T.lj);
    auto pflag = S->host.take<uint8_t>(pn), cg_pert = S->host.take<uint8_t>(any_pert ? ncg : 0);
    auto pa_iac = S->host.take<int32_t>(pn), pb_iac = S->host.take<int32_t>(pn);
    auto pa_q = S->host.take<double>(pn), pb_q = S->host.take<double>(pn), palj = S->host.take<double>(pn),
         pacrf = S->host.take<double>(pn);
    if (S->host.used > S->host.cap) return fail(GROMOS_AMD_EHIP, "layout staging overflow (internal)");
    // pass 2 (parallel over jobs): each job fills its own slices, indices
    // shifted by its atom offset
    auto fill = [&](int j) {
        const JobImage& im = *imgs[j];
        const Slice& o = sl[j];
        const int32_t A = (int32_t)o.atom;
        for (int c = 0; c < im.n_cg; ++c) {
            const int64_t g = o.cg + c;
            cg_start[g] = A + im.cg_start[c];
            cg_n[g] = im.cg_start[c + 1] - im.cg_start[c];
            cg_job[g] = j;
            if (any_pert) {
                uint8_t f = 0;
                if (im.has_pert)
                    for (int a = im.cg_start[c]; a < im.cg_start[c + 1]; ++a) f |= im.pflag[a];
                cg_pert[g] = f ? 1 : 0;
            }
        }
        const int64_t npad_j = sl[j + 1].atom - o.atom;
        for (int a = 0; a < im.n; ++a) {
            atom_job[A + a] = j;
            q[A + a] = im.q[a];
            iac[A + a] = im.iac[a];
            mass[A + a] = im.mass[a];
            inv_mass[A + a] = im.mass[a] > 0.0 ? 1.0 / im.mass[a] : 0.0;
        }
        for (int64_t a = im.n; a < npad_j; ++a) {
            atom_job[A + a] = -1;
            iac[A + a] = 0;
            q[A + a] = 0.0;
            mass[A + a] = 0.0;
            inv_mass[A + a] = 0.0;
        }
        for (int64_t a = 0; a < npad_j; ++a) {
            const int32_t la = (int32_t)std::min<int64_t>(a, im.n);
            excl_off[A + a] = (int32_t)(o.excl + im.excl_off[la]);
            o14_off[A + a] = (int32_t)(o.o14 + im.o14_off[la]);
            o14s_off[A + a] = (int32_t)(o.o14s + im.o14s_off[la]);
        }
        for (size_t k = 0; k < im.excl.size(); ++k) excl[o.excl + k] = A + im.excl[k];
        for (size_t k = 0; k < im.o14.size(); ++k) o14[o.o14 + k] = A + im.o14[k];
        for (size_t k = 0; k < im.o14s.size(); ++k) {
            const int32_t e = im.o14s[k];
            o14s[o.o14s + k] = (int32_t)((uint32_t)(A + (e & 0x7FFFFFFF)) | ((uint32_t)e & 0x80000000u));
        }
        if (any_pert)
            for (int64_t a = 0; a < npad_j; ++a) {
                const bool pa = im.has_pert && a < im.n && im.pflag[a];
                pflag[A + a] = pa ? 1 : 0;
                pa_iac[A + a] = pa ? im.pa_iac[a] : 0;
                pb_iac[A + a] = pa ? im.pb_iac[a] : 0;
                pa_q[A + a] = pa ? im.pa_q[a] : 0.0;
                pb_q[A + a] = pa ? im.pb_q[a] : 0.0;
                palj[A + a] = pa ? im.palj[a] : 0.0;
                pacrf[A + a] = pa ? im.pacrf[a] : 0.0;
            }
        if (im.P.use_settle)
            for (size_t k = 0; k < im.mol_first.size(); ++k) mol_first[o.molf + k] = A + im.mol_first[k];
        for (size_t k = 0; k < im.papairs.size(); k += 4) {
            papairs[o.pap + k] = A + im.papairs[k];
            papairs[o.pap + k + 1] = A + im.papairs[k + 1];
            papairs[o.pap + k + 2] = im.papairs[k + 2];
            papairs[o.pap + k + 3] = im.papairs[k + 3];
        }
        std::copy(im.lj.begin(), im.lj.end(), lj.data() + o.lj);
        for (int32_t a = 0; a < im.J.n_solute_atoms; ++a) sol_atom[o.sol + a] = A + a;
        for (size_t k = 0; k < im.bonded.atoms.size(); ++k) {
            const int4 e = im.bonded.atoms[k];
            bd_atoms[o.bd + k] = make_int4(e.x < 0 ? -1 : A + e.x, e.y < 0 ? -1 : A + e.y, e.z < 0 ? -1 : A + e.z,
                                           e.w < 0 ? -1 : A + e.w);
            bd_prm[o.bd + k] = im.bonded.prm[k];
            bd_kind[o.bd + k] = im.bonded.kind[k];
        }
    };
    {
        // a thread per 32k atoms, at most one per job and at most 16 (or the cores)
        const int64_t hw = std::max<int64_t>(1, std::min<int64_t>(16, (int64_t)std::thread::hardware_concurrency()));
        const int nt = (int)std::min<int64_t>(std::min<int64_t>(njobs, hw), std::max<int64_t>(1, npad / 32768));
        if (nt <= 1) {
            for (int j = 0; j < njobs; ++j) fill(j);
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; ++t)
                th.emplace_back([&, t]() {
                    for (int j = t; j < njobs; j += nt) fill(j);
                });
            for (auto& x : th) x.join();
        }
    }
    excl_off[npad] = (int32_t)T.excl;
    o14_off[npad] = (int32_t)T.o14;
    o14s_off[npad] = (int32_t)T.o14s;
    S->npad = (int32_t)npad;
    S->ncg = (int32_t)ncg;
    S->ncell = (int32_t)cell_off;
    S->nmol = (int32_t)mol_first.size();
    S->nrow_blocks = (int32_t)(align_up(npad * 4, BLOCK) / BLOCK);   // 4 lanes per atom in the pair kernel
    S->nsnb = std::max(1, I0.P.nsnb);
    S->use_settle = use_settle;
    S->do_virial = do_virial || any_baro;   // the pressure needs every step's virial
    S->twin = twin;
    S->any_thermo = any_thermo;
    S->any_com = any_com;
    S->any_baro = any_baro;
    S->any_solute_shake = any_solute_shake;
    S->any_solvent_pass = any_solvent_pass;
    S->job_step0.assign((size_t)njobs, 0);
    S->host_pressure.assign((size_t)njobs, 0.0);
    S->near_r = VV_NEAR_R;
    // the far-pair kernel's persistent form (GROMOS_AMD_VV_PERSIST=1): three
    // resident blocks of four waves per CU at its three waves per SIMD
    if (const char* env = std::getenv("GROMOS_AMD_VV_PERSIST"); env && std::atoi(env) == 1) {
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        S->vv_persist = std::max(1, cus) * 3;
    } else {
        S->vv_persist = 0;
    }
    if (const char* env = std::getenv("GROMOS_AMD_VV_NEAR_R")) S->near_r = std::max(0.0, std::atof(env));
    // LJ only between the solvent molecules' first atoms (SPC-like waters)
    S->vv_lj00 = 1;
    for (const JobDev& J : S->jobs)
        if (J.n_atoms > J.n_solute_atoms)
            for (int i = 0; i < J.aps; ++i)
                for (int jj = 0; jj < J.aps; ++jj)
                    if ((i || jj) && (J.sv_c6[i * 4 + jj] != 0.0 || J.sv_c12[i * 4 + jj] != 0.0)) S->vv_lj00 = 0;
    // neutral solvent molecules (the charges of one molecule sum to zero exactly)
    S->vv_neutral = 1;
    for (const JobImage* im : imgs)
        if (im->solvent) {
            const int32_t f = im->J.n_solute_atoms, a = im->J.aps;
            double sum = 0.0;
            for (int i = 0; i < a; ++i) sum += im->q[f + i];
            if (sum != 0.0) S->vv_neutral = 0;
        }
    // GROMOS_AMD_VV_PAIRWISE_RF=1: the per-atom-pair RF form everywhere (checks)
    if (const char* env = std::getenv("GROMOS_AMD_VV_PAIRWISE_RF"))
        if (std::atoi(env) == 1) S->vv_neutral = 0;
    S->cell_bits = 1;
    while ((1ll << S->cell_bits) < (int64_t)S->ncell + 1) S->cell_bits++;
    S->n_papairs = (int32_t)papairs.size() / 4;
    S->nsol = (int32_t)sol_atom.size();
    // FP32 records of the perturbed solute atoms for the unperturbed partners'
    // shares: charges, alphas, soft-core RF constants (perturbed.rs:160-190), types
    std::vector<float4> prec(2 * std::max<int64_t>(T.sol, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    for (int j = 0; j < njobs; ++j) {
        const JobImage& im = *imgs[j];
        if (!im.has_pert) continue;
        const JobDev& J = S->jobs[j];
        for (int32_t a = 0; a < J.n_solute_atoms; ++a) {
            if (!im.pflag[a]) continue;
            const int64_t w = sl[j].sol + a;
            const double cut2 = J.crf.cutoff_sq, acrf = im.pacrf[a];
            const double a_c2s = cut2 + acrf * J.lp.b_crfs_lambda2, b_c2s = cut2 + acrf * J.lp.a_crfs_lambda2;
            const double a_cut3i = 2.0 * J.crf.crf_2cut3i / std::sqrt(a_c2s * a_c2s * a_c2s);
            const double b_cut3i = 2.0 * J.crf.crf_2cut3i / std::sqrt(b_c2s * b_c2s * b_c2s);
            float ai, bi;
            std::memcpy(&ai, &im.pa_iac[a], 4);
            std::memcpy(&bi, &im.pb_iac[a], 4);
            prec[2 * w] = make_float4((float)im.pa_q[a], (float)im.pb_q[a], (float)im.palj[a], (float)acrf);
            prec[2 * w + 1] = make_float4((float)a_cut3i, (float)b_cut3i, ai, bi);
        }
    }
    S->has_pert = any_pert;
    // chunks of at most VV_CHUNK sorted groups, each within one job (a job's
    // groups keep its own range [cg_off, cg_off + n_cg) in sorted space)
    std::vector<int4> vv_chunk;
    for (int j = 0; j < njobs; ++j) {
        const int32_t g0 = S->jobs[j].cg_off, g1 = g0 + imgs[j]->n_cg;
        for (int32_t c = g0; c < g1; c += VV_CHUNK) vv_chunk.push_back(make_int4(c, std::min(c + VV_CHUNK, g1), j, 0));
    }
    S->nvv_chunk = (int32_t)vv_chunk.size();
    // per-atom index of the PERTATOMPAIR entries (both atoms of an entry),
    // entries of an atom in entry order
    PinnedArena::Span<int32_t> pap_off, pap;
    if (any_pert) {
        pap_off = S->host.take<int32_t>(np1);
        std::fill(pap_off.begin(), pap_off.end(), 0);
        for (int32_t e = 0; e < S->n_papairs; ++e) {
            int32_t i = papairs[4 * e], jj = papairs[4 * e + 1];
            if (i == jj) continue;   // zero distance: skipped by the reference (perturbed.rs:668-671)
            pap_off[i + 1]++;
            pap_off[jj + 1]++;
        }
        for (int64_t a = 0; a < npad; ++a) pap_off[a + 1] += pap_off[a];
        pap = S->host.take<int32_t>(pap_off[npad]);
        if (S->host.used > S->host.cap) return fail(GROMOS_AMD_EHIP, "layout staging overflow (internal)");
        std::vector<int32_t> fillp(pap_off.begin(), pap_off.end() - 1);
        for (int32_t e = 0; e < S->n_papairs; ++e) {
            int32_t i = papairs[4 * e], jj = papairs[4 * e + 1];
            if (i == jj) continue;
            pap[fillp[i]++] = (int32_t)((uint32_t)e | 0x80000000u);
            pap[fillp[jj]++] = e;
        }
    }

    // covalent terms: per solute atom (sol_atom order) its contributions in term order
    S->nbd = (int32_t)T.bd;
    PinnedArena::Span<int32_t> bd_aoff, bd_aref;
    if (T.bd > 0) {
        bd_aoff = S->host.take<int32_t>((size_t)T.sol + 1);
        std::fill(bd_aoff.begin(), bd_aoff.end(), 0);
        // global atom -> solute slot of its job
        auto slot = [&](int32_t a) -> int32_t {
            const int32_t j = atom_job[a];
            return S->jobs[j].sol_off + (a - S->jobs[j].atom_off);
        };
        int64_t nent = 0;
        for (int64_t k = 0; k < T.bd; ++k)
            for (int32_t v : {bd_atoms[k].x, bd_atoms[k].y, bd_atoms[k].z, bd_atoms[k].w})
                if (v >= 0) {
                    bd_aoff[slot(v) + 1]++;
                    nent++;
                }
        for (int64_t w = 0; w < T.sol; ++w) bd_aoff[w + 1] += bd_aoff[w];
        bd_aref = S->host.take<int32_t>((size_t)nent);
        if (S->host.used > S->host.cap) return fail(GROMOS_AMD_EHIP, "layout staging overflow (internal)");
        std::vector<int32_t> at(bd_aoff.begin(), bd_aoff.end() - 1);
        for (int64_t k = 0; k < T.bd; ++k) {
            const int32_t* ap = &bd_atoms[k].x;
            for (int s2 = 0; s2 < 4; ++s2)
                if (ap[s2] >= 0) bd_aref[at[slot(ap[s2])]++] = (int32_t)((k << 2) | s2);
        }
    }
    S->concat_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_concat).count();
    hipStream_t st = S->stream;
    int rc = 0;
    rc = rc ? rc : S->upload(S->d_jobs, S->jobs);
    rc = rc ? rc : S->upload(S->d_pg_start, S->host_pg);
    rc = rc ? rc : S->upload(S->d_sh_ij, S->host_sh_ij);
    rc = rc ? rc : S->upload(S->d_sh_len, S->host_sh_len);
    rc = rc ? rc : S->upload(S->d_lc_sdiag, S->host_lc_sdiag);
    rc = rc ? rc : S->upload(S->d_lc_cpl, S->host_lc_cpl);
    rc = rc ? rc : S->upload(S->d_lc_cidx, S->host_lc_cidx);
    rc = rc ? rc : S->upload(S->d_lc_coef, S->host_lc_coef);
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
    rc = rc ? rc : S->upload(S->d_pflag, pflag);
    rc = rc ? rc : S->upload(S->d_pa_iac, pa_iac);
    rc = rc ? rc : S->upload(S->d_pb_iac, pb_iac);
    rc = rc ? rc : S->upload(S->d_pa_q, pa_q);
    rc = rc ? rc : S->upload(S->d_pb_q, pb_q);
    rc = rc ? rc : S->upload(S->d_palj, palj);
    rc = rc ? rc : S->upload(S->d_pacrf, pacrf);
    rc = rc ? rc : S->upload(S->d_cg_pert, cg_pert);
    rc = rc ? rc : S->upload(S->d_sol_atom, sol_atom);
    rc = rc ? rc : S->upload(S->d_papairs, papairs);
    rc = rc ? rc : S->upload(S->d_pap_off, pap_off);
    rc = rc ? rc : S->upload(S->d_pap, pap);
    rc = rc ? rc : S->upload(S->d_vv_chunk, vv_chunk);
    rc = rc ? rc : S->upload(S->d_prec, prec);
    if (T.bd > 0) {
        rc = rc ? rc : S->upload(S->d_bd_atoms, bd_atoms);
        rc = rc ? rc : S->upload(S->d_bd_prm, bd_prm);
        rc = rc ? rc : S->upload(S->d_bd_kind, bd_kind);
        rc = rc ? rc : S->upload(S->d_bd_aoff, bd_aoff);
        rc = rc ? rc : S->upload(S->d_bd_aref, bd_aref);
        rc = rc ? rc : S->reserve(S->d_bd_tf, (size_t)T.bd * 12, true);
        rc = rc ? rc : S->reserve(S->d_bd_tev, (size_t)T.bd * BD_EV, true);
        rc = rc ? rc : S->reserve(S->d_bd_fb, (size_t)T.sol * 3, true);
    }
    {
        // the atom blocks of BLOCK atoms that hold solute atoms: in md steps the
        // excluded / 1-4 / self terms of solvent atoms are taken elsewhere
        std::vector<int32_t>& eb = S->h_excl_blocks;
        eb.clear();
        for (const JobDev& J : S->jobs)
            for (int32_t b = J.atom_off / BLOCK; b < (J.atom_off + J.n_solute_atoms + BLOCK - 1) / BLOCK; ++b)
                eb.push_back(b);
        S->n_excl_blocks = (int32_t)eb.size();
        rc = rc ? rc : S->upload(S->d_excl_blocks, eb);
    }
    if (rc) return rc;
    size_t n3 = (size_t)npad * 3, nc = (size_t)ncg, ncl = (size_t)S->ncell + 1, npa = (size_t)npad;
    for (int k = 0; k < 2; ++k) {
        if ((rc = S->reserve(S->d_x[k], n3, true))) return rc;
        if ((rc = S->reserve(S->d_v[k], n3, true))) return rc;
        HIPCHK(hipMemsetAsync(S->d_x[k], 0, n3 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(S->d_v[k], 0, n3 * sizeof(double), st));
    }
    if ((rc = S->reserve(S->d_f, n3, true)) || (rc = S->reserve(S->d_ref, 3 * nc, true)) ||
        (rc = S->reserve(S->d_s_ref, 3 * nc, true)) || (rc = S->reserve(S->d_key_in, nc, true)) ||
        (rc = S->reserve(S->d_val_in, nc, true)) || (rc = S->reserve(S->d_key_out, nc, true)) ||
        (rc = S->reserve(S->d_val_out, nc, true)) || (rc = S->reserve(S->d_cell_count, ncl, true)) ||
        (rc = S->reserve(S->d_cell_start, ncl, true)) || (rc = S->reserve(S->d_s_cg_orig, nc, true)) ||
        (rc = S->reserve(S->d_s_cg_n, nc + 1, true)) || (rc = S->reserve(S->d_s_cg_scan, nc + 1, true)) ||
        (rc = S->reserve(S->d_s_cg_start, nc, true)) || (rc = S->reserve(S->d_s2o, npa, true)) ||
        (rc = S->reserve(S->d_s_atom_cg, npa, true)) || (rc = S->reserve(S->d_o2s, npa, true)) ||
        (rc = S->reserve(S->d_s_type, npa, true)) || (rc = S->reserve(S->d_s_prec, npa, true)) || (rc = S->reserve(S->d_vv_cnt, nc + 1, true)) ||
        (rc = S->reserve(S->d_gen_cnt, nc + 1, true)) || (rc = S->reserve(S->d_vv_off, nc + 1, true)) ||
        (rc = S->reserve(S->d_gen_off, nc + 1, true)) || (rc = S->reserve(S->d_ctl, 6 + JOB_CTL * (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_gsol, nc, true)) ||
        (rc = S->reserve(S->d_pc_short, nc, true)) || (rc = S->reserve(S->d_pc_long, nc, true)) ||
        (rc = S->reserve(S->d_xq, npa, true)) || (rc = S->reserve(S->d_vv_cap, nc + 1, true)) ||
        (rc = S->reserve(S->d_gen_cap, nc + 1, true)) || (rc = S->reserve(S->d_prev_vv, nc, true)) ||
        (rc = S->reserve(S->d_prev_gen, nc, true)) ||
        (rc = S->reserve(S->d_f_s, n3, true)) || (rc = S->reserve(S->d_f_long_s, n3, true)) ||
        (rc = S->reserve(S->d_f_x, n3, true)) ||
        (rc = S->reserve(S->d_xf, npa, true)) || (rc = S->reserve(S->d_fi_s, n3, true)) ||
        (rc = S->reserve(S->d_fi_long_s, n3, true)) ||
        (rc = S->reserve(S->d_rows, (size_t)NSLOT * S->nrow_blocks * ROW, true)) ||
        (rc = S->reserve(S->d_settle_fail, (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_force_fail, (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_therm_scale, (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_nh_zeta, (size_t)GROMOS_AMD_NHC_MAX * njobs, true)) ||
        (rc = S->reserve(S->d_job_step0, (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_press, (size_t)10 * njobs, true)) ||
        (rc = S->reserve(S->d_box_fail, (size_t)njobs, true)) ||
        (rc = S->reserve(S->d_cg_cvir, any_baro ? 3 * nc : 1, true)) ||
        (rc = S->reserve(S->d_sh_skip, any_solute_shake ? 2 * (size_t)npad : 1, true)) ||
        (rc = S->reserve(S->d_lc_work, any_lincs_u ? 6 * S->host_sh_len.size() : 1, true)) ||
        (rc = S->reserve(S->d_nh_pend, (size_t)GROMOS_AMD_NHC_MAX * njobs, true)) ||
        (rc = S->reserve(S->d_vv_e, 11 * nc, true)) ||
        (rc = S->reserve(S->d_vv_part, (size_t)PART_W * VV_PARTS * njobs, true)) ||
        (rc = S->reserve(S->d_cg_ek, 3 * nc, true)) || (rc = S->reserve(S->d_step, 1)) ||
        (rc = S->reserve(S->d_vv_near, nc, true)) || (rc = S->reserve(S->d_fi_near, n3, true)) ||
        (rc = S->reserve(S->d_vv_sched, (size_t)NXCD_SCHED)) ||
        (rc = S->reserve(S->d_vv_en, 11 * nc, true)) ||
        (rc = S->reserve(S->d_step_done, 1)) ||
        (rc = S->reserve(S->d_sol_e, (size_t)2 * ROW * std::max<int32_t>(S->nsol, 1), true)))
        return rc;
    S->d_counters = S->d_ctl;
    S->d_overflow = (int32_t*)(S->d_ctl + 5);
    S->d_job_pairs = (int64_t*)(S->d_ctl + 6);
    if (S->h_ctl_n < 6 + JOB_CTL * (size_t)njobs) {
        if (S->h_ctl) HIPCHK(hipHostFree(S->h_ctl));
        S->h_ctl = nullptr;
        S->h_ctl_n = 0;
        HIPCHK(hipHostMalloc((void**)&S->h_ctl, (6 + JOB_CTL * (size_t)njobs) * sizeof(uint64_t), hipHostMallocDefault));
        S->h_ctl_n = 6 + JOB_CTL * (size_t)njobs;
    }
    // set once per layout, kept by every build: padding atoms' sorted slots map to
    // no atom; cell counts start at zero (each build's k_scatter counts them back
    // to zero for the next); sorted copies of padding atoms are zero
    HIPCHK(hipMemsetAsync(S->d_s2o, 0xFF, npa * sizeof(int32_t), st));
    HIPCHK(hipMemsetAsync(S->d_cell_count, 0, ncl * sizeof(int32_t), st));
    HIPCHK(hipMemsetAsync(S->d_xq, 0, npa * sizeof(double4), st));
    HIPCHK(hipMemsetAsync(S->d_xf, 0, npa * sizeof(float4), st));
    HIPCHK(hipMemsetAsync(S->d_f_long_s, 0, n3 * sizeof(double), st));
    HIPCHK(hipMemsetAsync(S->d_vv_near, 0, nc * sizeof(int32_t), st));
    // the persistent far-pair kernel's counters start at zero (each launch's last
    // wave zeroes them again)
    HIPCHK(hipMemsetAsync(S->d_vv_sched, 0, NXCD_SCHED * sizeof(unsigned), st));
    // Nose-Hoover chains start at zero (nosehoover_thermostat.rs:64-113); a
    // session puts continuing jobs' chains back after the layout
    HIPCHK(hipMemsetAsync(S->d_nh_zeta, 0, (size_t)GROMOS_AMD_NHC_MAX * njobs * sizeof(double), st));
    HIPCHK(hipMemsetAsync(S->d_press, 0, (size_t)10 * njobs * sizeof(double), st));
    HIPCHK(hipMemsetAsync(S->d_fi_near, 0, n3 * sizeof(long long), st));   // padding slots stay zero
    HIPCHK(hipMemsetAsync(S->d_step, 0, sizeof(int32_t), st));
    HIPCHK(hipMemsetAsync(S->d_step_done, 0, sizeof(uint32_t), st));
    HIPCHK(hipMemsetAsync(S->d_rows, 0, (size_t)NSLOT * S->nrow_blocks * ROW * sizeof(double), st));
    HIPCHK(hipMemsetAsync(S->d_fi_long_s, 0, n3 * sizeof(long long), st));
    HIPCHK(hipMemsetAsync(S->d_s_cg_n, 0, (nc + 1) * sizeof(int32_t), st));
    HIPCHK(hipMemsetAsync(S->d_vv_cnt, 0, (nc + 1) * sizeof(int64_t), st));
    HIPCHK(hipMemsetAsync(S->d_gen_cnt, 0, (nc + 1) * sizeof(int64_t), st));
    HIPCHK(hipMemsetAsync(S->d_vv_cap, 0, (nc + 1) * sizeof(int64_t), st));
    HIPCHK(hipMemsetAsync(S->d_gen_cap, 0, (nc + 1) * sizeof(int64_t), st));
    // scratch for the library scans
    size_t t2 = 0, t3 = 0, t4 = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, S->d_cell_count, S->d_cell_start, S->ncell + 1, st));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, S->d_s_cg_n, S->d_s_cg_scan, S->ncg + 1, st));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, S->d_vv_cnt, S->d_vv_off, S->ncg + 1, st));
    size_t tb = std::max(t2, std::max(t3, t4));
    if ((rc = S->reserve(S->d_temp, tb, true))) return rc;
    S->temp_bytes = S->bufs[(void*)&S->d_temp].bytes;
    HIPCHK(hipStreamSynchronize(st));
    S->cur = 0;
    S->therm_dirty = true;
    S->xq_valid = false;
    S->built = false;
    S->have_prev = false;
    S->ctl_pending = false;
    S->long_computed = false;
    S->steps_since_update = 0;
    return 0;
}

int upload_jobs(gromos_amd_system* S) {
    // the host copies of barostat boxes follow the device after every md call
    // (md_steps), so this upload keeps them
    HIPCHK(hipMemcpyAsync(S->d_jobs, S->jobs.data(), S->jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// StandardPairlistAlgorithm::update (pairlist.rs:155-174) on the device.
// Cell sort, then the neighbour lists.  After the first build the lists are
// filled in one pass into slots sized from each group's previous count; if a
// group outgrew its slot (flag) the lists are refilled with the exact counts
// that pass produced.
// ---------------------------------------------------------------------------
static int scan_offsets(gromos_amd_system* S, const int64_t* cnt_vv, const int64_t* cnt_gen) {
    size_t tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, cnt_vv, S->d_vv_off, S->ncg + 1, S->stream));
    tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, cnt_gen, S->d_gen_off, S->ncg + 1, S->stream));
    return 0;
}

// the control block's readback (enqueued after the search) -> host counters
static int parse_ctl(gromos_amd_system* S) {
    for (int c = 0; c < 5; ++c) S->host_counters[c] = S->h_ctl[c];
    for (size_t b = 0; b < JOB_CTL * (size_t)S->njobs; ++b) S->host_job_pairs[b] = (int64_t)S->h_ctl[6 + b];
    return 0;
}

static int read_ctl(gromos_amd_system* S) {
    const size_t nctl = 6 + JOB_CTL * (size_t)S->njobs;
    HIPCHK(hipMemcpyAsync(S->h_ctl, S->d_ctl, nctl * sizeof(uint64_t), hipMemcpyDeviceToHost, S->stream));
    HIPCHK(hipStreamSynchronize(S->stream));
    return parse_ctl(S);
}

// the pair kernels hold list offsets in 32 bits
int check_list_size(int64_t vv, int64_t gen) {
    if (vv >= MAX_LIST_ENTRIES || gen >= MAX_LIST_ENTRIES)
        return fail(GROMOS_AMD_ENOMEM,
                    "neighbour lists of %lld + %lld entries exceed the 32-bit list offsets (%lld); run fewer or "
                    "smaller jobs per batch", (long long)vv, (long long)gen, (long long)MAX_LIST_ENTRIES);
    return 0;
}

// the search's near radius: mixed precision evaluates the near solvent
// partners in FP64; the FP64 path lists every partner as near (one ascending run)
static double near_r2(const gromos_amd_system* S) {
    return S->precision == GROMOS_AMD_MIXED ? S->near_r * S->near_r : 1e300;
}

// lists with exact slots: the counts of the last search pass scanned into
// offsets, buffers sized to the totals, a second fill pass
static int refill_exact(gromos_amd_system* S) {
    hipStream_t st = S->stream;
    const int32_t ncg = S->ncg;
    if (int rc = scan_offsets(S, S->d_vv_cnt, S->d_gen_cnt)) return rc;
    int64_t tot[2];
    HIPCHK(hipMemcpyAsync(&tot[0], S->d_vv_off + ncg, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&tot[1], S->d_gen_off + ncg, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (int rc = read_ctl(S)) return rc;
    if (int rc = check_list_size(tot[0], tot[1])) return rc;
    if (int rc = S->reserve(S->d_vv_ent, (size_t)tot[0], true)) return rc;
    if (int rc = S->reserve(S->d_gen_ent, (size_t)tot[1], true)) return rc;
    SysPtrs p = S->ptrs();
    launch_caps(p, ncg, 1, st);
    // the refill lists the solvent groups with general partners again (the host
    // keeps the counts it read)
    HIPCHK(hipMemsetAsync(S->d_counters + 4, 0, sizeof(uint64_t), st));
    launch_nbl(p, ncg, S->ncell, true, near_r2(S), st);
    HIPCHK(hipGetLastError());
    return 0;
}

// bookkeeping once the host holds this build's counts
static void build_done(gromos_amd_system* S) {
    if (S->njobs == 1) {   // one job: its counts are the totals
        S->host_job_pairs[0] = (int64_t)S->host_counters[0];
        S->host_job_pairs[1] = (int64_t)S->host_counters[1];
        S->host_job_pairs[2] = (int64_t)S->host_counters[2];
    }
    S->prev_tot_vv = (int64_t)S->host_counters[2];
    S->prev_tot_gen = (int64_t)S->host_counters[3];
}

// deferred: the single-pass fill's readback is left in flight (S->ctl_pending);
// finish_pairlist waits for it, and refills if a group outgrew its slot.
int build_pairlist(gromos_amd_system* S, bool deferred) {
    hipStream_t st = S->stream;
    SysPtrs p = S->ptrs();
    int32_t ncg = S->ncg;
    const size_t nctl = 6 + JOB_CTL * (size_t)S->njobs;
    // reference points, cell keys and counts (the control block cleared here)
    launch_cg_ref(p, ncg, (int32_t)nctl, st);
    size_t tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, S->d_cell_count, S->d_cell_start, S->ncell + 1, st));
    launch_sort_cg(p, ncg, S->ncell, st);
    tb = S->temp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(S->d_temp, tb, S->d_s_cg_n, S->d_s_cg_scan, ncg + 1, st));
    // sorted maps and the sorted position copies (the gather of this order)
    launch_sorted_atoms(p, ncg, S->precision == GROMOS_AMD_MIXED ? 1 : 0, st);
    S->xq_valid = true;
    S->built = true;
    S->rebuilds++;
    S->steps_since_update = 0;
    if (S->have_prev) {
        // GROMOS_AMD_TIGHT_SLOTS=1 (tests): slots exactly the previous counts, so
        // that the overflow and refill path runs at most builds
        const char* tenv = std::getenv("GROMOS_AMD_TIGHT_SLOTS");
        launch_caps(p, ncg, tenv && std::atoi(tenv) == 1 ? 2 : 0, st);
        if (int rc = scan_offsets(S, S->d_vv_cap, S->d_gen_cap)) return rc;
        int64_t bv = S->prev_tot_vv + S->prev_tot_vv / 8 + 16 * (int64_t)ncg + 64;
        int64_t bg = S->prev_tot_gen + S->prev_tot_gen / 8 + 16 * (int64_t)ncg + 64;
        if (int rc = check_list_size(bv, bg)) return rc;
        if (int rc = S->reserve(S->d_vv_ent, (size_t)bv, true)) return rc;
        if (int rc = S->reserve(S->d_gen_ent, (size_t)bg, true)) return rc;
        p = S->ptrs();
        launch_nbl(p, ncg, S->ncell, true, near_r2(S), st);
        if (S->njobs > 1) launch_job_pairs(p, S->njobs, st);
        HIPCHK(hipMemcpyAsync(S->h_ctl, S->d_ctl, nctl * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(S->ctl_ev, st));
        HIPCHK(hipGetLastError());
        S->ctl_pending = true;
        if (deferred) return 0;
        bool refilled = false;
        return finish_pairlist(S, &refilled);
    }
    // first build: count pass, then exact slots
    launch_nbl(p, ncg, S->ncell, false, near_r2(S), st);
    if (S->njobs > 1) launch_job_pairs(p, S->njobs, st);
    if (int rc = refill_exact(S)) return rc;
    build_done(S);
    S->have_prev = true;
    return 0;
}

int finish_pairlist(gromos_amd_system* S, bool* refilled) {
    *refilled = false;
    if (!S->ctl_pending) return 0;
    S->ctl_pending = false;
    HIPCHK(hipEventSynchronize(S->ctl_ev));
    parse_ctl(S);
    if ((int32_t)S->h_ctl[5]) {   // overflow flag (low word): the fill pass left exact counts
        S->refills++;
        // kernels on the side stream may still read the lists the refill replaces
        if (S->side) HIPCHK(hipStreamSynchronize(S->side));
        if (int rc = refill_exact(S)) return rc;
        *refilled = true;
    }
    build_done(S);
    return 0;
}

// Autotuner of the mixed-precision solvent pair kernel: lanes per molecule (4, 8,
// 16) chosen by timing two launches of each on the batch itself, the first time a
// batch shape (molecule size, LJ layout, virial, log2 of the group count) is met
// in this process; GROMOS_AMD_VV_LANES forces a value.  That kernel adds fixed-
// point groups of entries, so every choice gives the same bits and the step
// that tunes is an ordinary step.
static std::mutex g_tune_mu;
static std::map<int64_t, int32_t> g_tuned;

static int tune_vv(gromos_amd_system* S, const SysPtrs& p, int32_t aps, int32_t mixed) {
    int32_t bucket = 0;
    while ((1 << bucket) < S->ncg) ++bucket;
    const int32_t form = (S->vv_lj00 ? VV_LJ00 : 0) | (S->vv_neutral ? VV_NEUTRAL : 0);
    int64_t key = (((((int64_t)bucket * 8 + aps) * 2 + mixed) * 4 + form) * 2 + S->do_virial);
    if (S->vv_lanes && S->tuned_key == key) return 0;
    if (const char* env = std::getenv("GROMOS_AMD_VV_LANES")) {
        int v = std::atoi(env);
        S->vv_lanes = (v == 4 || v == 16) ? v : 8;
        S->tuned_key = key;
        return 0;
    }
    if (!mixed) {
        // the FP64 path sums each lane's pairs in FP64, so its bits depend on the
        // lane count: keep one fixed count (results independent of the batch)
        S->vv_lanes = 8;
        S->tuned_key = key;
        return 0;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            S->vv_lanes = it->second;
            S->tuned_key = key;
            return 0;
        }
    }
    // each candidate timed twice (order 4 8 16 16 8 4), the faster launch counts:
    // a single cold launch can be slowed by cache and clock state
    const int32_t cand[3] = {4, 8, 16};
    const int order[6] = {0, 1, 2, 2, 1, 0};
    hipEvent_t ev[7];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(ev[0], S->stream));
    for (int c = 0; c < 6; ++c) {
        launch_vv(p, S->ncg, aps, form, S->do_virial, mixed, cand[order[c]], S->stream);
        HIPCHK(hipEventRecord(ev[c + 1], S->stream));
    }
    HIPCHK(hipEventSynchronize(ev[6]));
    float best_of[3] = {1e30f, 1e30f, 1e30f};
    for (int c = 0; c < 6; ++c) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[c], ev[c + 1]));
        best_of[order[c]] = std::min(best_of[order[c]], ms);
    }
    int32_t best = 8;
    float best_ms = 1e30f;
    for (int c = 0; c < 3; ++c)
        if (best_of[c] < best_ms) {
            best_ms = best_of[c];
            best = cand[c];
        }
    for (auto& e : ev) (void)hipEventDestroy(e);
    HIPCHK(hipGetLastError());
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = best;
    }
    S->vv_lanes = best;
    S->tuned_key = key;
    return 0;
}

// atoms per solvent molecule of the batch (all jobs with solvent share it)
static int32_t solvent_aps(const gromos_amd_system* S) {
    for (const JobDev& J : S->jobs)
        if (J.n_atoms > J.n_solute_atoms) return J.aps;
    return S->jobs[0].aps > 0 ? S->jobs[0].aps : 3;
}

// pair interactions of one step (which = 0: short range, 1: long range), only
// over the jobs that advance in it: a job past its stop step is held until the
// next batch change and its pairs are not evaluated
static int64_t active_pairs(const gromos_amd_system* S, int which) {
    bool all = true;
    for (const JobDev& J : S->jobs) all = all && S->md_step < J.stop;
    if (all) return (int64_t)S->host_counters[which];
    int64_t n = 0;
    for (int32_t b = 0; b < S->njobs; ++b)
        if (S->md_step < S->jobs[b].stop) n += S->host_job_pairs[JOB_CTL * (size_t)b + which];
    return n;
}

// Forcefield::apply nonbonded part (forcefield.rs:226-1060); energy partials
// into the rows and records.  assemble: total forces into d_f (the md step
// assembles them itself).
int forcefield(gromos_amd_system* S, bool assemble) {
    hipStream_t st = S->stream;
    if (!S->built) {
        if (int rc = build_pairlist(S)) return rc;
    }
    bool rebuild = S->steps_since_update >= S->nsnb;   // pairlist.rs:105-107
    if (rebuild) {
        // the counts' readback stays in flight while the first pair kernels are
        // enqueued (finish_pairlist below)
        if (int rc = build_pairlist(S, true)) return rc;
    }
    S->steps_since_update++;
    SysPtrs p = S->ptrs();
    int32_t npad = S->npad;
    const int32_t mixed = S->precision == GROMOS_AMD_MIXED;
    // sorted position copies: written by the previous step's integrator or by the
    // pair-list build (new order), unless the state was set from outside
    if (!S->xq_valid) launch_gather(p, npad, mixed, st);
    S->xq_valid = false;
    gromos_amd_system::TimedLaunch tl{};
    if (S->timing) {
        tl.a0 = S->next_event();
        tl.a1 = S->next_event();
        if (mixed) {
            tl.n0 = S->next_event();
            tl.n1 = S->next_event();
        }
    }
    hipEvent_t e0 = tl.a0, e1 = tl.a1, n0 = tl.n0, n1 = tl.n1;
    const int32_t aps = solvent_aps(S);
    if (int rc = tune_vv(S, p, aps, mixed)) return rc;
    const int32_t form = (S->vv_lj00 ? VV_LJ00 : 0) | (S->vv_neutral ? VV_NEUTRAL : 0);
    // mixed md step: the near solvent pairs (FP64) and the solute atoms' pass
    // (general pairs, then excluded / 1-4 terms, covalent terms) are independent
    // of the far solvent pair kernel (separate outputs, inputs ready at the fork),
    // so they overlap it on the side stream
    const bool split = mixed && !assemble && S->nsol > 0 && !S->no_side;
    auto launch_first = [&](hipEvent_t a, hipEvent_t b, hipEvent_t na, hipEvent_t nb) -> int {
        if (split) {
            if (na) {
                tl.f = S->next_event();
                HIPCHK(hipEventRecord(tl.f, st));
            }
            HIPCHK(hipEventRecord(S->ev_fork, st));
            HIPCHK(hipStreamWaitEvent(S->side, S->ev_fork, 0));
            launch_vv_near(p, aps, form, S->do_virial, S->side, na, nb);
            launch_gen(p, npad, false, S->do_virial, mixed, 0, S->side, GEN_SOLUTE);
            launch_excl14(p, npad, 0, S->d_excl_blocks, S->n_excl_blocks, S->side);
            launch_bonded(p.bd, p.x_cur, S->side);
            HIPCHK(hipEventRecord(S->ev_join, S->side));
        }
        launch_vv(p, S->ncg, aps, form, S->do_virial, mixed, S->vv_lanes, st, a, b);
        if (mixed && !split) launch_vv_near(p, aps, form, S->do_virial, st, na, nb);
        return 0;
    };
    if (int rc = launch_first(e0, e1, n0, n1)) return rc;
    bool refilled = false;
    if (int rc = finish_pairlist(S, &refilled)) return rc;
    if (refilled) {
        // a group outgrew its slot: the kernels above ran on truncated lists;
        // they assign their outputs, so running them again on the refilled lists
        // replaces everything they wrote
        p = S->ptrs();
        if (int rc = launch_first(nullptr, nullptr, nullptr, nullptr)) return rc;
    }
    // the timed solvent pair launch's work: the advancing jobs' solvent-solvent
    // molecule pairs (each listed from both sides) times the atom pairs of one
    if (e0) {
        S->timed.push_back(tl);
        S->timed_pairs += active_pairs(S, 2) / 2 * (int64_t)aps * aps;
    }
    const int32_t ngsol = (int32_t)S->host_counters[4];
    launch_gen(p, npad, false, S->do_virial, mixed, ngsol, st, split ? GEN_SOLVENT : GEN_ALL);
    S->pair_count += active_pairs(S, 0);
    if (S->twin && (rebuild || !S->long_computed)) {
        launch_gen(p, npad, true, 1, mixed, ngsol, st);
        S->long_computed = true;
        S->pair_count += active_pairs(S, 1);
    }
    if (assemble) launch_excl14(p, npad, 1, nullptr, 0, st);
    else if (!split) launch_excl14(p, npad, 0, S->d_excl_blocks, S->n_excl_blocks, st);
    // covalent terms (forcefield.rs:386-400): per-atom sums in bd.fb, added last
    if (!split) launch_bonded(p.bd, p.x_cur, st);
    if (assemble) {
        launch_assemble(p, npad, S->twin ? 1 : 0, mixed, st);
        launch_bonded_add(p.bd, p.f, st);
    }
    if (split) HIPCHK(hipStreamWaitEvent(st, S->ev_join, 0));
    HIPCHK(hipGetLastError());
    return 0;
}

// One plain md step as a graph: the pair kernels, the correction passes, the
// integrator and the reductions of a step without a rebuild, captured once per
// state-buffer parity and replayed while every kernel argument (the batch's
// device pointers and shapes) is unchanged; the step index comes from the
// device counter.  No memset or host copy is captured.
static int launch_step_graph(gromos_amd_system* S) {
    hipStream_t st = S->stream;
    const SysPtrs p = S->ptrs();
    const int32_t mixed = S->precision == GROMOS_AMD_MIXED;
    const int32_t aps = solvent_aps(S);
    const int32_t form = (S->vv_lj00 ? VV_LJ00 : 0) | (S->vv_neutral ? VV_NEUTRAL : 0);
    const int64_t shape[11] = {S->ncg, S->npad, S->njobs, S->nsol, S->vv_lanes, form, S->do_virial, S->twin ? 1 : 0,
                               mixed, aps, (int64_t)S->host_counters[4]};
    auto& g = S->graphs[S->cur];
    if (!g.exec || std::memcmp(&g.key, &p, sizeof p) != 0 || std::memcmp(g.shape, shape, sizeof shape) != 0) {
        if (g.exec) HIPCHK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launch_vv(p, S->ncg, aps, form, S->do_virial, mixed, S->vv_lanes, st);
        if (mixed) launch_vv_near(p, aps, form, S->do_virial, st);
        launch_gen(p, S->npad, false, S->do_virial, mixed, (int32_t)S->host_counters[4], st);
        launch_excl14(p, S->npad, 0, S->d_excl_blocks, S->n_excl_blocks, st);
        launch_bonded(p.bd, p.x_cur, st);
        launch_integrate(p, S->ncg, S->twin ? 1 : 0, 1, mixed, st);
        launch_reduce(p, S->npad, S->twin ? 1 : 0, 1, 1, S->njobs, st);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) return fail(GROMOS_AMD_EHIP, "step graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            return fail(GROMOS_AMD_EHIP, "step graph instantiate: %s", hipGetErrorString(e));
        }
        g.key = p;
        std::memcpy(g.shape, shape, sizeof shape);
        S->graph_captures++;
    }
    HIPCHK(hipGraphLaunch(g.exec, st));
    S->graph_launches++;
    return 0;
}

int md_steps(gromos_amd_system* S, int32_t nsteps, std::vector<double>* rec, std::vector<int32_t>* failed_out,
             std::vector<uint8_t>* force_out) {
    if (nsteps <= 0) return 0;
    int64_t need = (int64_t)nsteps * S->njobs * REC;
    if (int rc = S->reserve(S->d_records, (size_t)need, true)) return rc;
    hipStream_t st = S->stream;
    std::vector<int32_t> big(S->njobs, INT32_MAX);
    HIPCHK(hipMemcpyAsync(S->d_settle_fail, big.data(), big.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S->d_force_fail, big.data(), big.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S->d_box_fail, big.data(), big.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    size_t ev0 = S->step_ev.size();
    auto step_event = [&]() -> int {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        S->step_ev.push_back(e);
        HIPCHK(hipEventRecord(e, st));
        return 0;
    };
    if (S->step_timing)
        if (int rc = step_event()) return rc;
    const int32_t mixed = S->precision == GROMOS_AMD_MIXED;
    HIPCHK(hipMemsetAsync(S->d_step, 0, sizeof(int32_t), st));   // never inside a captured graph
    if (S->any_com)
        HIPCHK(hipMemcpyAsync(S->d_job_step0, S->job_step0.data(), S->job_step0.size() * sizeof(int64_t),
                              hipMemcpyHostToDevice, st));
    // graph replay of the plain steps is opt-in (GROMOS_AMD_GRAPHS=1): on this
    // runtime it measured 1.5 % slower than separate launches (water_32768)
    const char* genv = std::getenv("GROMOS_AMD_GRAPHS");
    S->use_graphs = genv && std::atoi(genv) == 1;
    const char* senv = std::getenv("GROMOS_AMD_SIDE_STREAM");
    S->no_side = senv && std::atoi(senv) == 0;
    for (int32_t s = 0; s < nsteps; ++s) {
        // a plain step (no rebuild now or next, sorted copies current, no pending
        // thermostat rescan, no per-kernel timing events): one captured graph
        const bool plain = S->use_graphs && S->built && S->steps_since_update + 1 < S->nsnb && S->xq_valid &&
                           !(S->any_thermo && S->therm_dirty) && !S->timing && !(S->twin && !S->long_computed) &&
                           !S->any_com && !S->any_baro && !S->any_solute_shake && !S->any_solvent_pass &&
                           S->vv_lanes > 0;
        if (plain) {
            S->steps_since_update++;
            S->md_step = s;
            S->pair_count += active_pairs(S, 0);
            if (int rc = launch_step_graph(S)) return rc;
            S->cur = 1 - S->cur;   // exchange_state (configuration.rs:426-428)
            if (S->step_timing)
                if (int rc = step_event()) return rc;
            continue;
        }
        S->md_step = s;
        if (int rc = forcefield(S, false)) return rc;
        SysPtrs p = S->ptrs();
        if (S->any_thermo && S->therm_dirty) {
            launch_therm(p, S->njobs, st);
            S->therm_dirty = false;
        }
        // centre-of-mass motion removal (first in the reference's step, md.rs:1095-1097):
        // on the velocities entering the step, after the bath read their E_kin
        if (S->any_com) launch_com(p, S->njobs, st);
        if (S->any_baro) launch_press_ke(p, S->njobs, st);
        // the integrator leaves the sorted copies for the next step unless that
        // step rebuilds the list (pairlist.rs:105-107)
        const int32_t write_xq = S->steps_since_update < S->nsnb ? 1 : 0;
        launch_integrate(p, S->ncg, S->twin ? 1 : 0, write_xq, mixed, st);
        if (S->any_solute_shake) launch_shake_solute(p, S->njobs, st);
        if (S->any_solvent_pass) launch_solvent_constraints(p, S->ncg, st);
        launch_reduce(p, S->npad, S->twin ? 1 : 0, 1, 1, S->njobs, st);
        if (S->any_baro) launch_baro(p, S->njobs, st);
        S->cur = 1 - S->cur;   // exchange_state (configuration.rs:426-428)
        // the barostat scaled the new positions after the integrator copied them
        // (and the constraints after the integrator moved positions after that copy)
        S->xq_valid = write_xq != 0 && !S->any_baro && !S->any_solute_shake && !S->any_solvent_pass;
        if (S->step_timing)
            if (int rc = step_event()) return rc;
    }
    S->md_step = 0;
    S->last_rec_row = nsteps - 1;
    for (int32_t b = 0; b < S->njobs; ++b) S->job_step0[b] += std::min<int64_t>(nsteps, S->jobs[b].stop);
    HIPCHK(hipGetLastError());
    std::vector<int32_t> failed(S->njobs), ffail(S->njobs), bfail(S->njobs, INT32_MAX);
    HIPCHK(hipMemcpyAsync(failed.data(), S->d_settle_fail, failed.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ffail.data(), S->d_force_fail, ffail.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    std::vector<JobDev> jd;
    std::vector<double> press;
    if (S->any_baro) {
        // the barostat's boxes and pressures, kept on the host side too
        jd.resize(S->jobs.size());
        press.resize((size_t)10 * S->njobs);
        HIPCHK(hipMemcpyAsync(bfail.data(), S->d_box_fail, bfail.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(jd.data(), S->d_jobs, jd.size() * sizeof(JobDev), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(press.data(), S->d_press, press.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (rec) {
        rec->resize((size_t)need);
        HIPCHK(hipMemcpyAsync(rec->data(), S->d_records, (size_t)need * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (S->step_timing && S->step_ev.size() > ev0) {
        if (int rc = collect_step_times(S)) return rc;
    }
    if (int rc = S->collect_timing()) return rc;
    for (size_t b = 0; b < jd.size(); ++b) {
        S->jobs[b].box = jd[b].box;
        S->jobs[b].use_cg_grouped = jd[b].use_cg_grouped;
        S->host_pressure[b] = press[10 * b + 9];
    }
    // a force failure at step s leaves that step's forces undefined: the job
    // stops there like a SETTLE failure; so does a box the cell grid no longer fits
    std::vector<uint8_t> kind(S->njobs, 0);
    for (int j = 0; j < S->njobs; ++j) {
        if (ffail[j] <= failed[j] && ffail[j] != INT32_MAX) {
            failed[j] = ffail[j];
            kind[j] = 1;
        }
        if (bfail[j] < failed[j]) {
            failed[j] = bfail[j];
            kind[j] = 2;
        }
    }
    if (failed_out) {
        *failed_out = failed;
        if (force_out) *force_out = kind;
        return 0;
    }
    for (int j = 0; j < S->njobs; ++j)
        if (failed[j] != INT32_MAX) {
            if (kind[j] == 2)
                return fail(GROMOS_AMD_EBOX, "the barostat shrank the box below the pair search's cells (job %d, step %d)",
                            j, failed[j]);
            if (kind[j])
                return fail(GROMOS_AMD_EFORCE, "non-finite or out-of-range force (|F| >= %.3g kJ/mol/nm; job %d, step %d)",
                            FORCE_LIMIT, j, failed[j]);
            return fail(GROMOS_AMD_ESETTLE, "constraints failed: SETTLE sin(angle) > 1.0 or SHAKE unconverged after 1000 sweeps (job %d, step %d)", j, failed[j]);
        }
    return 0;
}

int collect_step_times(gromos_amd_system* S) {
    for (size_t k = 0; k + 1 < S->step_ev.size(); ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, S->step_ev[k], S->step_ev[k + 1]));
        S->step_ms.push_back(ms);
    }
    for (hipEvent_t e : S->step_ev) (void)hipEventDestroy(e);
    S->step_ev.clear();
    return 0;
}

}  // namespace gamd

// ---------------------------------------------------------------------------
// gromos_amd_system members
// ---------------------------------------------------------------------------
gromos_amd_system::~gromos_amd_system() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_ctl) (void)hipHostFree(h_ctl);
    for (auto& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& kv : bufs)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : step_ev) (void)hipEventDestroy(e);
    if (side) (void)hipStreamSynchronize(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ctl_ev) (void)hipEventDestroy(ctl_ev);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
}

hipEvent_t gromos_amd_system::next_event() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
}

// the solvent pair work of each timed step: k_vv (FP64) or k_vv_mx followed by
// the near pairs' k_vv_nr on one stream (the two durations summed), or, where
// k_vv_nr runs on the side stream beside k_vv_mx, the span from the first of the
// two starting to the last finishing (overlapping durations are not added)
int gromos_amd_system::collect_timing() {
    if (timed.empty()) {
        ev_used = 0;
        return 0;
    }
    HIPCHK(hipStreamSynchronize(stream));
    if (side) HIPCHK(hipStreamSynchronize(side));
    auto el = [](hipEvent_t a, hipEvent_t b, float& ms) { return hipEventElapsedTime(&ms, a, b); };
    for (const TimedLaunch& t : timed) {
        float ms = 0.f;
        if (t.f) {
            float a0 = 0.f, a1 = 0.f, n0 = 0.f, n1 = 0.f;
            HIPCHK(el(t.f, t.a0, a0));
            HIPCHK(el(t.f, t.a1, a1));
            HIPCHK(el(t.f, t.n0, n0));
            HIPCHK(el(t.f, t.n1, n1));
            ms = std::max(a1, n1) - std::min(a0, n0);
        } else {
            HIPCHK(el(t.a0, t.a1, ms));
            if (t.n0) {
                float mn = 0.f;
                HIPCHK(el(t.n0, t.n1, mn));
                ms += mn;
            }
        }
        timed_ms += ms;
        timed_launches++;
    }
    timed.clear();
    ev_used = 0;
    return 0;
}

SysPtrs gromos_amd_system::ptrs() const {
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
    p.gsol = d_gsol;
    p.pc_short = d_pc_short;
    p.pc_long = d_pc_long;
    p.job_pairs = d_job_pairs;
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
    p.rows = d_rows;
    p.vv_e = d_vv_e;
    p.fi_near = d_fi_near;
    p.vv_en = precision == GROMOS_AMD_MIXED ? d_vv_en : nullptr;
    p.vv_part = d_vv_part;
    p.cg_ek = d_cg_ek;
    p.sol_atom = d_sol_atom;
    p.nsol = nsol;
    p.vv_chunk = d_vv_chunk;
    p.nvv_chunk = nvv_chunk;
    p.vv_sched = d_vv_sched;
    p.vv_persist = vv_persist;
    p.step = d_step;
    p.prec = d_prec;
    p.s_prec = d_s_prec;
    p.step_done = d_step_done;
    p.sol_e = d_sol_e;
    p.vv_virial = do_virial;
    p.records = d_records;
    p.settle_fail = d_settle_fail;
    p.force_fail = d_force_fail;
    p.therm_scale = d_therm_scale;
    p.nh_zeta = d_nh_zeta;
    p.nh_pend = d_nh_pend;
    p.job_step0 = d_job_step0;
    p.press = d_press;
    p.pg_start = d_pg_start;
    p.box_fail = d_box_fail;
    p.cg_cvir = d_cg_cvir;
    p.sh_ij = d_sh_ij;
    p.sh_len = d_sh_len;
    p.sh_skip = d_sh_skip;
    p.lc_sdiag = d_lc_sdiag;
    p.lc_cpl = d_lc_cpl;
    p.lc_cidx = d_lc_cidx;
    p.lc_coef = d_lc_coef;
    p.lc_work = d_lc_work;
    p.pflag = d_pflag;
    p.pa_iac = d_pa_iac;
    p.pb_iac = d_pb_iac;
    p.pa_q = d_pa_q;
    p.pb_q = d_pb_q;
    p.palj = d_palj;
    p.pacrf = d_pacrf;
    p.cg_pert = d_cg_pert;
    p.papairs = d_papairs;
    p.pap_off = d_pap_off;
    p.pap = d_pap;
    p.o2s = d_o2s;
    p.vv_near = d_vv_near;
    p.bd = gamd::BondedDev{};
    if (nbd > 0) {
        p.bd.atoms = d_bd_atoms;
        p.bd.prm = d_bd_prm;
        p.bd.kind = d_bd_kind;
        p.bd.tf = d_bd_tf;
        p.bd.tev = d_bd_tev;
        p.bd.aoff = d_bd_aoff;
        p.bd.aref = d_bd_aref;
        p.bd.amap = d_sol_atom;
        p.bd.fb = d_bd_fb;
        p.bd.nterms = nbd;
        p.bd.natoms = nsol;
    }
    return p;
}
