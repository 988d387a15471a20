"""Copies of the oracle with one plausible mistake planted each.

Each mutant names a text edit of oracle/qpad_oracle.c (a dropped term, a
wrong sign, a transposed component, a wrong axis factor or boundary fold)
and the pins of tests/oracle_pins.py that must fail on it.  A mistake in a
matrix is named only for the matrix pin: the source pins recover the
right-hand side with the library's own matrix, so they test the sources
given the matrix.  The copies are compiled with the oracle's own flags into
a temporary directory; nothing here touches the built oracle.  Mistakes are
planted only in this CPU code.
"""
import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "oracle", "qpad_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared"]

# name: (original text, planted text, pins that must fail)
MUTANTS = {
    "bplus_index_m": ("double t = (double)(m + 1) / j;", "double t = (double)(m) / j;",
                      ["matrices"]),
    "bminus_axis_relax": ("if (m == 1) { a[0] = 0.0; b[0] = -4.0 - relax; c[0] = 4.0; }",
                          "if (m == 1) { a[0] = 0.0; b[0] = -4.0; c[0] = 4.0; }",
                          ["matrices"]),
    "bt_log_fold_sign": ("if (m == 0) coef = 1.0 + 1.0 / (jmax * log(jmax * dr));",
                         "if (m == 0) coef = 1.0 - 1.0 / (jmax * log(jmax * dr));", ["matrices"]),
    "open_fold_m_plus_1": ("else coef = 1.0 - (double)m / jmax;\n            break;\n        case ORA_FK_BT:",
                           "else coef = 1.0 - (double)(m + 1) / jmax;\n            break;\n        case ORA_FK_BT:",
                           ["matrices"]),
    "ez_coupling_sign": ("bre[i] = idrh * (JRE(1, i + 1) - JRE(1, i - 1)) + ir * JRE(1, i) - dm * ir * JIM(2, i);",
                         "bre[i] = idrh * (JRE(1, i + 1) - JRE(1, i - 1)) + ir * JRE(1, i) + dm * ir * JIM(2, i);",
                         ["ez"]),
    "ez_gauss_edge": ("* ((double)nr - 2.5);", "* ((double)nr - 2.0);", ["ez"]),
    "ez_m1_node2_drop_term": (
        "bre[2] = idr * (JRE(1, 3) - JRE(1, 2)) + ir * JRE(1, 2) - dm * ir * JIM(2, 2);",
        "bre[2] = idr * (JRE(1, 3) - JRE(1, 2)) - dm * ir * JIM(2, 2);", ["ez"]),
    "bz_coupling_sign": ("bim[i] = -idrh * (JIM(2, i + 1) - JIM(2, i - 1)) - ir * JIM(2, i) + dm * ir * JRE(1, i);",
                         "bim[i] = -idrh * (JIM(2, i + 1) - JIM(2, i - 1)) - ir * JIM(2, i) - dm * ir * JRE(1, i);",
                         ["bz"]),
    "bz_axis_factor": ("bre[1] = -2.0 * idr * JRE(2, 2);", "bre[1] = -1.0 * idr * JRE(2, 2);", ["bz"]),
    "eperp_coupling_sign": ("FA(e, 1, r, i) = -FA(b, 0, r, i) + ir * dm * FA(psi, 0, q, i);",
                            "FA(e, 1, r, i) = -FA(b, 0, r, i) - ir * dm * FA(psi, 0, q, i);", ["eperp"]),
    "eperp_m1_axis_half": ("FA(e, 0, r, 1) = FA(b, 1, r, 1) - idr * FA(psi, 0, r, 2);",
                           "FA(e, 0, r, 1) = FA(b, 1, r, 1) - idrh * FA(psi, 0, r, 2);", ["eperp"]),
    "djdxi_transposed_flux": (
        "DR(1, i) = AR(1, i) - idrh * (MR(1, i + 1) - MR(1, i - 1)) - ir * MR(1, i) + dm * ir * MI(2, i);",
        "DR(1, i) = AR(1, i) - idrh * (MR(1, i + 1) - MR(1, i - 1)) - ir * MR(1, i) + dm * ir * MI(3, i);",
        ["djdxi"]),
    "djdxi_m2_node2_drop_term": (
        "DR(1, 2) = AR(1, 2) - idr * (MR(1, 3) - MR(1, 2)) - ir * MR(1, 2) + dm * ir * MI(2, 2);",
        "DR(1, 2) = AR(1, 2) - idr * (MR(1, 3) - MR(1, 2)) + dm * ir * MI(2, 2);", ["djdxi"]),
    "bt_iter_relax_sign": ("b1r[i] = -F1R(2, i) - F3R(1, i) * rx;", "b1r[i] = -F1R(2, i) + F3R(1, i) * rx;",
                           ["bt_iter"]),
    "bt_iter_recombination": ("FA(bf, 1, kr, i) = 0.5 * (b1i[i] - b2i[i]);",
                              "FA(bf, 1, kr, i) = 0.5 * (b2i[i] - b1i[i]);", ["bt_iter"]),
    "bt_iter_jz_coupling_sign": ("s1r = -F1R(2, i) + dm * F2I(3, i) * ir;",
                                 "s1r = -F1R(2, i) - dm * F2I(3, i) * ir;", ["bt_iter"]),
    "bt_beam_br_sign": ("FA(bf, 0, r, i) = -ir * dm * bim[i];", "FA(bf, 0, r, i) = ir * dm * bim[i];",
                        ["bt_beam"]),
    "amj_axis_m1_factor": ("FA(cu, 0, k, 1) = 8.0 * FA(cu, 0, k, 1); FA(cu, 1, k, 1) = 8.0 * FA(cu, 1, k, 1);",
                           "FA(cu, 0, k, 1) = 4.0 * FA(cu, 0, k, 1); FA(cu, 1, k, 1) = 8.0 * FA(cu, 1, k, 1);",
                           ["amjdeposit"]),
    "amj_axis_m2_flux_zeroed": ("for (int c = 0; c < 3; c++) FA(amu, c, k, 1) = 8.0 * FA(amu, c, k, 1);",
                                "for (int c = 0; c < 3; c++) FA(amu, c, k, 1) = 0.0 * FA(amu, c, k, 1);",
                                ["amjdeposit"]),
    # text shared by the robust and the standard deposit: planted in both
    "amj_flux_missing_ipsi": ("u2[0] = u[0] * u[0] * ipsi;", "u2[0] = u[0] * u[0];",
                              ["amjdeposit", "amjdeposit_std", "amjdeposit_pgc"], 3),
    "amj_phase_not_conjugated": ("double ph0r = pc, ph0i = -ps;", "double ph0r = pc, ph0i = ps;",
                                 ["amjdeposit", "amjdeposit_std", "amjdeposit_pgc"], 3),
    "std_normalisation_sign": ("double qtmh2 = qtmh / (1.0 - sp->qbm * sp->psi[p]);",
                               "double qtmh2 = qtmh / (1.0 + sp->qbm * sp->psi[p]);", ["amjdeposit_std"]),
    "std_weight_uses_gamma": ("double ipsi = 1.0 / (1.0 - sp->qbm * sp->psi[p]);",
                              "double ipsi = 1.0 / (sp->gamma[p] - u[2]);", ["amjdeposit_std"]),
    "std_second_kick_old_gamma": ("qtmh1 = qtmh2 * gam;\n        for (int c = 0; c < 3; c++) u[c] = ut[c] + ep[c] * qtmh1;",
                                  "for (int c = 0; c < 3; c++) u[c] = ut[c] + ep[c] * qtmh1;",
                                  ["amjdeposit_std"]),
    "std_push_own_gamma": ("double qtmh2 = qtmh1 * sp->gamma[p];",
                           "double qtmh2 = qtmh1 * sqrt(1.0 + pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);",
                           ["push_std"]),
    "interp_psi_weights_swapped": ("interp_vec(s, psi, 1, ix, w0, w1, pc, ps, &sp->psi[p]);",
                                   "interp_vec(s, psi, 1, ix, w1, w0, pc, ps, &sp->psi[p]);", ["interp_psi"]),
    "beam_xi_drift": ("bm->x3[p] = bm->x3[p] - bm->p3[p] * dt_gam + bm->dt;",
                      "bm->x3[p] = bm->x3[p] - bm->p3[p] * dt_gam;", ["beam_push"]),
    "beam_force_sign": ("wp[0] = ep[0] - bp[1];\n        wp[1] = ep[1] + bp[0];\n        wp[2] = ep[2];\n"
                        "        bm->p1[p]",
                        "wp[0] = ep[0] + bp[1];\n        wp[1] = ep[1] + bp[0];\n        wp[2] = ep[2];\n"
                        "        bm->p1[p]", ["beam_push"]),
    "beam_rotation": ("double t = ep[0] * cc - ep[1] * ss; double u = ep[0] * ss + ep[1] * cc;",
                      "double t = ep[0] * cc + ep[1] * ss; double u = ep[0] * ss + ep[1] * cc;", ["beam_push"]),
    "inject_cell_edge_lattice": ("double rn = ((double)i1 - 0.5) / (double)c->ppc1 + (double)(i - 1);",
                                 "double rn = ((double)i1) / (double)c->ppc1 + (double)(i - 1);", ["inject"]),
    "beam_quiet_charge": ("bm->q[i] = 0.5 * bm->q[i];\n            bm->q[n + i] = bm->q[i];\n"
                          "            bm->id[n + i] = bm->id[i] + (int32_t)c->total_num;",
                          "bm->q[i] = 1.0 * bm->q[i];\n            bm->q[n + i] = bm->q[i];\n"
                          "            bm->id[n + i] = bm->id[i] + (int32_t)c->total_num;", ["beam_inject"]),
    "lattice_quiet_charge": ("bm->q[i] = 0.5 * bm->q[i];\n            bm->q[n + i] = bm->q[i];\n"
                             "            bm->id[n + i] = bm->id[i] + (int32_t)n;",
                             "bm->q[i] = 1.0 * bm->q[i];\n            bm->q[n + i] = bm->q[i];\n"
                             "            bm->id[n + i] = bm->id[i] + (int32_t)n;", ["beam_lattice"]),
    "beam_sigma_swapped": ("xt[d] = g[d] * c->sigma[d] + c->center[d];",
                           "xt[d] = g[d] * c->sigma[2 - d] + c->center[d];", ["beam_inject"]),
    "lattice_rn_cell_edge": ("double rn = ((double)i1 - 0.5) / (double)c->ppc[0] + (double)(i - 1);",
                             "double rn = ((double)i1) / (double)c->ppc[0] + (double)(i - 1);",
                             ["beam_lattice"]),
    "lattice_gauss_sigma_squared": ("double t = (x - mu) / sigma;\n        return exp(-0.5 * (t * t));",
                                    "double t = (x - mu) / sigma;\n        return exp(-1.0 * (t * t));",
                                    ["beam_lattice"]),
    "laser_grad_phi_sign": ("FA(gr, 1, kr, i) = -ir * m * LZ(h, j, ki, i);",
                            "FA(gr, 1, kr, i) = ir * m * LZ(h, j, ki, i);", ["laser_grad"]),
    "laser_grad_xi_stencil": ("- 4.0 * LZ(h, j - 1, 0, i) + LZ(h, j - 2, 0, i));",
                              "- 4.0 * LZ(h, j - 1, 0, i) - LZ(h, j - 2, 0, i));", ["laser_grad"]),
    "laser_grad_axis_on_node1": ("const int ie = nr + 1;", "const int ie = 1;", ["laser_grad"]),
    "laser_grad_axis_factor": ("FA(gr, 0, kr, ie) = 2.0 * idrh * LZ(h, j, kr, 2);",
                               "FA(gr, 0, kr, ie) = idrh * LZ(h, j, kr, 2);", ["laser_grad"]),
    "pgc_quiver_factor": ("double gam_corr = 0.5 * sp->qbm * sp->qbm * (apr * apr + api * api);",
                          "double gam_corr = sp->qbm * sp->qbm * (apr * apr + api * api);", ["amjdeposit_pgc"]),
    "pgc_ponderomotive_xi_sign": ("ep[2] = ep[2] + tmp * (apr * aprg[2] + api * apig[2]);",
                                  "ep[2] = ep[2] - tmp * (apr * aprg[2] + api * apig[2]);",
                                  ["amjdeposit_pgc", "push_pgc"], 2),
    "pgc_drops_imaginary_envelope": ("ep[0] = ep[0] - tmp * (apr * aprg[0] + api * apig[0]);",
                                     "ep[0] = ep[0] - tmp * (apr * aprg[0]);", ["amjdeposit_pgc", "push_pgc"], 2),
    "pgc_robust_second_kick_initial_u": ("qtmh_e = std_ ? qtmh_b * gam : qtmh * gam / (gam - ut[2]);",
                                         "qtmh_e = std_ ? qtmh_b * gam : qtmh * gam / (gam - u0[2]);",
                                         ["amjdeposit_pgc"]),
    "pgc_std_kick_without_gamma": ("qtmh_b = qtmh / (1.0 - sp->qbm * sp->psi[p]);\n            qtmh_e = qtmh_b * gam;",
                                   "qtmh_b = qtmh / (1.0 - sp->qbm * sp->psi[p]);\n            qtmh_e = qtmh_b;",
                                   ["amjdeposit_pgc"]),
    "pgc_push_half_step_ahead": ("gam_corr = gam_corr + qbm2_hf * (apr + 0.25 * tmp) * tmp;",
                                 "gam_corr = gam_corr + qbm2_hf * (apr + 0.5 * tmp) * tmp;", ["push_pgc"]),
    "pgc_push_force_without_gamma": ("double tmp = 0.5 * sp->qbm / sp->gamma[p];",
                                     "double tmp = 0.5 * sp->qbm;", ["push_pgc"]),
    "chi_psi_sign": ("double phr = -1.0 * sp->qbm * sp->q[p] / (1.0 - sp->qbm * sp->psi[p]), phi = 0.0;",
                     "double phr = -1.0 * sp->qbm * sp->q[p] / (1.0 + sp->qbm * sp->psi[p]), phi = 0.0;",
                     ["deposit_chi"]),
    "chi_axis_factor_8": ("const double corr = (12.0 * ppc * ppc) / (1.0 + 2.0 * ppc * ppc);",
                          "const double corr = 8.0;", ["deposit_chi"]),
    "chi_uses_charge_not_q_over_m": ("double phr = -1.0 * sp->qbm * sp->q[p]",
                                     "double phr = -1.0 * sp->qm * sp->q[p]", ["deposit_chi"]),
    "adk_row_digit": ("{2.7226733893691e21, 2742.1316798375965, 1.0000920088118899}",
                      "{2.7226733893691e21, 2724.1316798375965, 1.0000920088118899}", ["adk_table"]),
    "ionize_first_order_only": ("double cons = mi[in] * w_ion[0] * dt * (1.0 + 0.5 * w_ion[0] * dt);",
                                "double cons = mi[in] * w_ion[0] * dt;", ["neutral_ionize"]),
    "ionize_no_overshoot_cap": ("            if (cons > mi[in]) {\n                shoot = 1;\n                cons = mi[in];",
                                "            if (cons > 2.0 * mi[in]) {\n                shoot = 1;\n                cons = mi[in];",
                                ["neutral_ionize"]),
    "ionize_reads_imaginary_part": ("const int kr = kre(m), kq = kre(m);", "const int kr = kre(m), kq = kim(m);",
                                    ["neutral_ionize"]),
    "ionize_total_truncated": ("(double)(int)(mi[ii] * (double)ppc_tot / (double)L + 0.5);",
                               "(double)(int)(mi[ii] * (double)ppc_tot / (double)L);", ["neutral_ionize"]),
    "add_lattice_cell_edge": ("const double rn = (double)(j - 1) + ((double)i - 0.5) / (double)ppc_add;",
                              "const double rn = (double)(j - 1) + ((double)i - 1.0) / (double)ppc_add;",
                              ["neutral_add"]),
    "add_charge_without_levels": ("const double coef = (double)L * copysign(1.0, sp->cfg.qm)",
                                  "const double coef = 1.0 * copysign(1.0, sp->cfg.qm)", ["neutral_add"]),
    "ion_charge_same_sign": ("sp->add_q[sp->n_add] = -sp->q[p];", "sp->add_q[sp->n_add] = sp->q[p];",
                             ["neutral_add"]),
    "update_bound_skips_one": ("while (i < sp->npp - 1) {\n        double pos = sqrt(sp->x1[i]",
                               "while (i < sp->npp - 2) {\n        double pos = sqrt(sp->x1[i]",
                               ["update_bound"]),
    "subcyc_step_floor": ("n = (int)ceil(exp_fac / exp_fac_max);", "n = (int)floor(exp_fac / exp_fac_max);",
                          ["subcyc_step"]),
    "clamp_without_sqrt": ("            sc = sqrt(sc);\n            sp->p1[p] = sp->p1[p] * sc;",
                           "            sp->p1[p] = sp->p1[p] * sc;", ["clamp_exp_fac"]),
    "exp_fac_from_zero": ("    double m = 1.0;", "    double m = 0.0;", ["exp_fac_max"]),
    "subcyc_push_full_step": ("slice_push(s, dsub, s->sc_fac_clamped);", "slice_push(s, s->g.dxi, s->sc_fac_clamped);",
                              ["subcyc_slice"]),
    "subcyc_without_clamp": ("slice_push(s, dsub, s->sc_fac_clamped);", "slice_push(s, dsub, 0.0);",
                             ["subcyc_slice"]),
    "subcyc_ionize_with_substep": (
        "            ora_neutral_ionize(s, sp, e);\n            ora_neutral_add(s, sp);",
        "            const double dt_keep = sp->dt;\n            sp->dt = dt;\n"
        "            ora_neutral_ionize(s, sp, e);\n            sp->dt = dt_keep;\n            ora_neutral_add(s, sp);",
        ["subcyc_slice"]),
    "subcyc_without_predictor": ("        slice_chi_cbq(s, j);\n        slice_predictor(s);\n        s->last_nsub = n;",
                                 "        slice_chi_cbq(s, j);\n        s->last_nsub = n;", ["subcyc_slice"]),
}


def source_of(name):
    """the oracle's source with mutant `name` planted: its text must occur as
    often as the mutant says (once, unless a fourth entry gives the count of
    places that share it), and every occurrence is replaced"""
    text = open(SRC).read()
    old, new = MUTANTS[name][:2]
    count = MUTANTS[name][3] if len(MUTANTS[name]) > 3 else 1
    assert text.count(old) == count, f"mutant {name}: the original text must occur exactly {count} time(s)"
    return text.replace(old, new)


def build_all(outdir, names=None):
    """compile every mutant (in parallel) into outdir; returns {name: path}"""
    names = sorted(MUTANTS) if names is None else names
    inc = os.path.join(HERE, "..", "oracle")

    def one(name):
        src = os.path.join(outdir, f"{name}.c")
        with open(src, "w") as f:
            f.write(source_of(name))
        lib = os.path.join(outdir, f"lib_{name}.so")
        subprocess.run(["gcc"] + CFLAGS + ["-I", inc, "-o", lib, src, "-lm"], check=True)
        return name, lib

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        return dict(ex.map(one, names))
