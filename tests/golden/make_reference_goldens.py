"""Writes reference_goldens.json: the known-answer data that the reference's own
tests hold for the hot path, transcribed as plain numbers (inputs and expected
outputs).  Each entry names the reference test file and lines it comes from.

Complex numbers are stored as [re, im] pairs; tensors as flat lists in the
reference's storage (column-major) order.  Run:  python tests/golden/make_reference_goldens.py
"""
import json
import os


def c(re, im=0.0):
    return [float(re), float(im)]


boundary_12 = [c(2, 0), c(0, 4), c(1, 0), c(0, 5), c(1, 0), c(0, 5), c(1, 0), c(1, 0),
               c(3, 0), c(0, 6), c(0, 2), c(-2, 0)]

goldens = {
    "contract_sos_left_non_trivial": {
        "source": "tests/boundaries.cpp:164-215",
        "dims": {"bl": 2, "br": 4, "cl": 3, "cr": 2, "d": 1},
        "L": boundary_12,
        "psi": [c(1, 0), c(1, 0), c(-1, 0), c(0, 2), c(1, 0), c(-1, 0), c(2, 0), c(-2, 0)],
        "idx": [1, 2, 3, 1],
        "mats": [c(1, 0), c(0, 1)],
        "expected": [c(-8, 1), c(4, 5), c(-10, -7), c(6, 0), c(-4, 5), c(8, 1), c(-14, 1), c(6, 0),
                     c(2, -7), c(-10, 1), c(16, -5), c(-6, 0), c(10, 16), c(10, -16), c(-10, 32), c(0, -12),
                     c(3, 9), c(3, -9), c(-3, 18), c(0, -12), c(1, -1), c(1, 1), c(-1, -2), c(0, 4),
                     c(0, 6), c(0, -6), c(0, 12), c(0, -12), c(6, -14), c(-10, 6), c(18, -16), c(-8, 16)],
    },
    "contract_sos_right_non_trivial": {
        "source": "tests/boundaries.cpp:250-303",
        "dims": {"bl": 4, "br": 2, "cl": 2, "cr": 3, "d": 1},
        "R": boundary_12,
        "psi": [c(1, 0), c(1, 0), c(1, 0), c(-1, 0), c(-1, 0), c(2, 0), c(0, 2), c(-2, 0)],
        "idx": [2, 1, 1, 3],
        "mats": [c(1, 0), c(0, 1)],
        "expected": [c(-8, 1), c(4, 5), c(-10, -7), c(18, 8), c(-4, 5), c(8, 1), c(-14, 1), c(18, -8),
                     c(2, -7), c(-10, 1), c(16, -5), c(-18, 16), c(-2, -8), c(-2, 8), c(2, -16), c(0, 20),
                     c(3, 9), c(3, -9), c(-3, 18), c(0, -24), c(1, -1), c(1, 1), c(-1, -2), c(0, 0),
                     c(0, 6), c(0, -6), c(0, 12), c(0, -12), c(-10, -6), c(6, 14), c(-14, -24), c(24, 24)],
    },
    "contract_sos_left_trivial_operator_dimension_2": {
        "source": "tests/boundaries.cpp:148-163",
        "dims": {"bl": 1, "br": 1, "cl": 1, "cr": 2, "d": 1},
        "L": [c(1)], "psi": [c(1)], "idx": [1, 2], "mats": [c(0, 1)],
        "expected": [c(0), c(0, 1)],
    },
    "contract_sos_right_trivial_operator_dimension_2": {
        "source": "tests/boundaries.cpp:234-249",
        "dims": {"bl": 1, "br": 1, "cl": 2, "cr": 1, "d": 1},
        "R": [c(1)], "psi": [c(1)], "idx": [2, 1], "mats": [c(0, 1)],
        "expected": [c(0), c(0, 1)],
    },
    "contract_expectation_boundaries_non_trivial": {
        "source": "tests/boundaries.cpp:96-129",
        "dims": {"b": 3, "c": 2},
        "L": [c(1, 2), c(2, 3), c(3, 4), c(4, 5), c(5, 6), c(6, 7), c(7, 8), c(8, 9), c(9, 10),
              c(1, 9), c(2, 8), c(3, 7), c(4, 6), c(5, 5), c(6, 4), c(7, 3), c(8, 2), c(9, 1)],
        "R": [c(1, 1), c(2, 2), c(3, 3), c(6, 6), c(5, 5), c(4, 4), c(7, 7), c(8, 8), c(9, 9),
              c(1, -1), c(2, -2), c(3, -3), c(6, -6), c(5, -5), c(4, -4), c(7, -7), c(8, -8), c(9, -9)],
        "expected": c(405, 495),
    },
    "compute_expectation_trivial": {
        "source": "tests/boundaries.cpp:21-84",
        "cases": [
            {"psi": [c(1)], "idx": [1, 1], "mats": [c(1)], "expected": c(1)},
            {"psi": [c(1)], "idx": [1, 1], "mats": [c(0, 1)], "expected": c(0, 1)},
            {"psi": [c(1), c(0)], "idx": [1, 1], "mats": [c(1), c(0), c(0), c(-1)], "expected": c(1)},
            {"psi": [c(0, 1), c(0)], "idx": [1, 1], "mats": [c(1), c(0), c(0), c(-1)], "expected": c(1)},
            {"psi": [c(0), c(1)], "idx": [1, 1], "mats": [c(1), c(0), c(0), c(-1)], "expected": c(-1)},
            {"psi": [c(0), c(0, 1)], "idx": [1, 1], "mats": [c(1), c(0), c(0), c(-1)], "expected": c(-1)},
        ],
    },
    "contract_vs_left_non_trivial": {
        "source": "tests/boundaries.cpp:342-383",
        "dims": {"blo": 3, "bro": 2, "bln": 2, "brn": 5, "d": 1},
        "L": [c(2, 0), c(0, 4), c(1, 0), c(0, 5), c(3, 0), c(0, 6)],
        "proj": [c(1, 0), c(0, -1), c(0, 1), c(1, 0), c(1, 0), c(0, -1)],
        "psi": [c(1, 0), c(1, 0), c(1, 0), c(-1, 0), c(-1, 0), c(2, 0), c(0, 2), c(-2, 0), c(1, -2), c(2, 0)],
        "expected": [c(0, 10), c(-6, 0), c(4, -6), c(5, 4), c(-4, 6), c(15, 15), c(-15, 9), c(-9, -21), c(21, 0), c(9, 21)],
    },
    "contract_vs_right_non_trivial": {
        "source": "tests/boundaries.cpp:422-465",
        "dims": {"blo": 3, "bro": 2, "bln": 2, "brn": 5, "d": 1},
        "R": [c(2, 0), c(1, 0), c(3, 0), c(1, 1), c(1, 2), c(0, 6), c(0, 4), c(0, 5), c(1, 1), c(2, 1)],
        "proj": [c(1, 0), c(0, -1), c(0, 1), c(1, 0), c(1, 0), c(0, -1)],
        "psi": [c(1, 0), c(1, 0), c(1, 0), c(-1, 0), c(-1, 0), c(2, 0), c(0, 2), c(-2, 0), c(1, -2), c(2, 0)],
        "expected": [c(5, -2), c(-3, -4), c(4, 5), c(26, 5), c(22, -15), c(-11, -2)],
    },
    "optimizer_external_field_one_site": {
        "source": "tests/optimizer.cpp:22-76",
        "note": "operator diag(1,...,1,-1) on one site with trivial boundaries: eigenvalue -1, eigenvector e_d, 0 iterations",
        "physical_dimensions": [1, 2, 3, 4],
    },
    "optimizer_modes": {
        "source": "tests/optimizer.cpp:77-136",
        "note": "one site, d = 4, operator diag(a,b,c,d); eigenvector is e_4",
        "cases": [
            {"which": "SR", "diag": [4, 0, -1, -2], "eigenvalue": -2},
            {"which": "LR", "diag": [-4, 0, 1, 2], "eigenvalue": 2},
            {"which": "LM", "diag": [2, 0, 1, -4], "eigenvalue": -4},
            {"which": "LM", "diag": [-2, 0, -1, 4], "eigenvalue": 4},
        ],
    },
    "tfim_sweep_until_converged": {
        "source": "tests/chain.cpp:286-305",
        "note": "reference TFIM H = sum Z_i + coupling sum X_i X_{i+1}; fixed bond dimension; relative tolerance 1e-7",
        "cases": [
            {"sites": 2, "coupling": 0.1, "chi": 2, "energy": -2.00249843},
            {"sites": 2, "coupling": 1.0, "chi": 2, "energy": -2.23606797},
            {"sites": 4, "coupling": 0.1, "chi": 2, "energy": -4.00750155},
            {"sites": 4, "coupling": 1.0, "chi": 4, "energy": -4.75877048},
            {"sites": 10, "coupling": 0.1, "chi": 2, "energy": -10.02251095},
            {"sites": 10, "coupling": 1.0, "chi": 6, "energy": -12.38148999},
        ],
    },
    "tfim_optimize_chain": {
        "source": "tests/chain.cpp:390-429",
        "note": "|E| of the reference TFIM ground state, relative tolerance 1e-10",
        "cases": [
            {"sites": 10, "coupling": 0.1, "abs_energy": 10.0225109571},
            {"sites": 10, "coupling": 1.0, "abs_energy": 12.3814899997},
        ],
    },
    "tfim_four_levels": {
        "source": "tests/stress.cpp:107-131",
        "note": "lowest four levels of the reference TFIM, absolute tolerance 1e-7",
        "cases": [
            {"sites": 6, "coupling": 0.1, "levels": [-6.012504691, -4.1912659256, -4.13264494449, -4.0501210912]},
            {"sites": 10, "coupling": 0.1, "levels": [-10.022510957, -8.2137057257, -8.18819723717, -8.1485537719]},
            {"sites": 10, "coupling": 0.5, "levels": [-10.569659557, -9.5030059614, -9.32268792732, -9.0714705801]},
            {"sites": 10, "coupling": 2.0, "levels": [-19.531007915, -19.5280782081, -17.3076728844, -17.3047431766]},
        ],
    },
    "external_field_sweep": {
        "source": "tests/chain.cpp:250-284",
        "note": "external field diag(-1,1,...,1) on every site: one sweep gives E = -N (relative 1e-7)",
        "cases": [[2, [2]], [3, [2]], [4, [2, 4]], [5, [2, 4]], [10, [2, 4, 8, 17]]],
    },
}

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")
with open(out, "w") as f:
    json.dump(goldens, f, indent=1)
print("wrote", out)
