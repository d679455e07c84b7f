"""Writes cook_membrane_q2.json: the inline mesh and the precomputed values of the reference's
Q2 user-element test (tests/test_user_elem_impl_diff_and_adaptive_load_step.py:29-57 mesh,
:64-66 Dirichlet/Neumann selection, :71-88 elements and traction, :122-124 material, :151-159
values).  The file is read as text and only its numeric array literals are taken (nothing of it
is imported or run).  Run in a checkout with the reference at /root/reference."""
import ast
import json
import os
import re

SRC = "/root/reference/tests/test_user_elem_impl_diff_and_adaptive_load_step.py"


def array_literal(text, name):
    m = re.search(name + r"\s*=\s*jnp\.array\(\s*(\[.*?\])\s*\)", text, re.S)
    return ast.literal_eval(m.group(1))


def main():
    text = open(SRC).read()
    values = re.findall(r"jnp\.array\(([-0-9.e+]+), dtype=jnp\.float64\)", text)
    names = ["dofs_dot_dofs", "d_dE", "d_dnu", "d2_dE2", "d2_dE_dnu", "d2_dnu_dE", "d2_dnu2"]
    fixture = {
        "source": "tests/test_user_elem_impl_diff_and_adaptive_load_step.py",
        "x_nodes": array_literal(text, "x_nodes"),
        "elements": array_literal(text, "elements"),
        "surface_elements": array_literal(text, "surface_elements"),
        "pts": [[0.0, 0.0], [48.0, 44.0], [48.0, 60.0], [0.0, 44.0]],
        "youngs_modulus": 100.0, "poisson_ratio": 0.3, "q_0": 20.0, "newton_tol": 1e-8,
        "gauss_order": 4,
        "golden": dict(zip(names, [float(v) for v in values])),
        "golden_rtol": 1e-5, "golden_atol": 1e-8,
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cook_membrane_q2.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", out, len(fixture["x_nodes"]), "nodes", len(fixture["elements"]), "elements",
          fixture["golden"])


if __name__ == "__main__":
    main()
