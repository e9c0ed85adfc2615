"""Pins the oracle's BP loop (bp_iterate and orc_perform_inference in
oracle/oracle.c) with the hand-computed cases of bp_loop_cases.py, and shows
that each case catches the planted mistakes its docstring names: every
mutation is compiled into a variant of the oracle on the CPU
(oracle.build_variant) and the case must fail on it."""
import pytest

import oracle
from bp_loop_cases import CHECKS, DIMS, K_UNIT, PRIOR, RES, SIG, camera


def _oracle_maker(library):
    def make(cams):
        o = oracle.Oracle(K_UNIT, 1, 1, RES, PRIOR, DIMS, [SIG, 0, 0], [0, 0, 0], lookup_n=3, activation=1,
                          library=library)
        for x0, d in cams:
            o.add_camera(*camera(x0, d))
        return o
    return make


# planted mistakes in oracle.c's BP loop, and the checks that must catch each
MUTANTS = {
    "S_not_divided_by_C": ([("double a = clampd(s / C);", "double a = clampd(s);")],
                           ["one_camera_one_iteration"]),
    "belief_update_without_minus_Mc": ([("c->B[v] = (c->B[v] + m) - k->M[v];", "c->B[v] = c->B[v] + m;")],
                                       ["fixed_point_of_one_camera"]),
    "gate_dropped": ([("        }\n        if (C > CUM_LEN_GATE) {", "        }\n        if (C > 0.0) {")],
                     ["cum_length_gate"]),
    "jacobi_order": ([("clampd(1.0 / (1.0 + exp(c->B[vox] - k->M[vox])))",
                       "clampd(1.0 / (1.0 + exp((log(c->A[vox]) - log(1.0 - c->A[vox])) - k->M[vox])))")],
                     ["gauss_seidel"]),
    "no_exclusion_of_own_message": ([("clampd(1.0 / (1.0 + exp(c->B[vox] - k->M[vox])))",
                                      "clampd(1.0 / (1.0 + exp(c->B[vox])))")],
                                    ["fixed_point_of_one_camera"]),
    "messages_kept_across_calls": ([("    for (int i = 0; i < c->ncams; ++i) memset(c->cams[i].M, 0, sizeof(double) * (size_t)nv);",
                                     "")],
                                   ["second_call_resets_messages"]),
    "logit_as_log": ([("double m = log(a) - log(1.0 - a);", "double m = log(a);")],
                     ["one_camera_one_iteration", "gauss_seidel"]),
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_bp_loop_by_hand(name):
    err = CHECKS[name](_oracle_maker(oracle.lib()))
    assert err < 1e-9, f"{name}: {err:.3e}"


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_is_caught(mutant, tmp_path):
    repl, catchers = MUTANTS[mutant]
    make = _oracle_maker(oracle.build_variant(repl, str(tmp_path), name=mutant))
    for name in catchers:
        err = CHECKS[name](make)
        assert err > 1e-3, f"{name} does not catch {mutant} ({err:.3e})"
