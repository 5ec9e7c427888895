"""Writes known_answers.json: the reference's own known-answer vectors.

Every value below is transcribed from HybridSmoothers.jl's test-suite (v0.1.4);
the file:line of each is stored with it.  Inputs are described, not stored as
matrices: the 1D Poisson matrix spdiagm(2, -1, -1) of test/test_matrices.jl:5.

    python tests/golden/make_known_answers.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

FIXTURES = {
    "n9_poisson": {
        "source": "test/test_preconditioners.jl:110-126 (GPU mirror test/gpu/gpu.jl:113-129)",
        "matrix": "poisson_test_matrix(9)",
        "x": [0, 1, 2, 3, 4, 5, 6, 7, 8],
        "partsize": 2,
        "eta": 1.5,
        "y_forward": [0, 0.5, 1.0, 2.0, 2.0, 3.5, 3.0, 5.0, 4.0],
        "y_backward": [0.25, 0.5, 1.75, 1.5, 3.25, 2.5, 4.75, 3.5, 4.0],
        "y_symmetric": [0.25, 0.5, 2.0, 2.0, 3.75, 3.5, 5.5, 5.0, 4.0],
        "D_DL1": [2, 2, 2, 2, 2, 2, 2, 2, 2],
        "SLbuffer": [-1, -1, -1, -1],
        "SUbuffer": [-1, -1, -1, -1],
        "chunks": [1, 2, 3, 4, 5, 6, 7, 8],
    },
    "eta_cases": {
        "source": "test/test_preconditioners.jl:278-317",
        "matrix": "poisson_test_matrix(9)",
        "partsize": 2,
        "chunks": 2,
        "sweep": "forward",
        "cases": [
            {"eta": 2.0, "D_DL1": [2, 2, 2, 2, 2, 2, 2, 2, 2], "SLbuffer": [-1, -1, -1, -1]},
            {"eta": 3.0, "D_DL1": [2.0, 2.5, 2.5, 2.5, 2.5, 2.5, 2.5, 2.5, 2.5],
             "SLbuffer": [-1, -1, -1, -1]},
            {"eta": 1.0, "D_DL1": [2, 2, 2, 2, 2, 2, 2, 2, 2]},
        ],
        "y_forward_eta2": [0, 0.5, 1.0, 2.0, 2.0, 3.5, 3.0, 5.0, 4.0],
    },
    "nonsym_csc": {
        "source": "test/test_preconditioners.jl:320-353",
        "matrix": "poisson_test_matrix(9) with A[1,8] = A[2,8] = -1 (1-based)",
        "extra_entries_1based": [[1, 8, -1.0], [2, 8, -1.0]],
        "x": [0, 1, 2, 3, 4, 5, 6, 7, 8],
        "partsize": 2,
        "chunks": 2,
        "y_forward": [0, 1.0 / 3.0, 1.0, 2.0, 2.0, 3.5, 3.0, 5.0, 4.0],
        "y_backward": [1.0 / 6.0, 1.0 / 3.0, 1.75, 1.5, 3.25, 2.5, 4.75, 3.5, 4.0],
        "D_DL1": [2, 3, 2, 2, 2, 2, 2, 2, 2],
        "SLbuffer": [-1, -1, -1, -1],
        "SUbuffer": [-1, -1, -1, -1],
    },
    "partsize3": {
        "source": "test/test_preconditioners.jl:355-364",
        "matrix": "poisson_test_matrix(9)",
        "partsize": 3,
        "chunks": 2,
        "sweep": "forward",
        "D_DL1": [2, 2, 2, 2, 2, 2, 2, 2, 2],
        "SLbuffer": [-1, 0, -1, -1, 0, -1, -1, 0, -1],
    },
    "l1_rowsum": {
        "source": "test/test_chebyshev.jl:15-19",
        "matrix": "cheb_poisson_1d(5)",
        "l1": [3.0, 4.0, 4.0, 4.0, 3.0],
    },
    "cheb1_astar_table": {
        "source": "src/chebyshev.jl:208-210 (three-digit table of D'Ambra et al. Fig. 3)",
        "k": [1, 2, 3, 4, 5, 6, 7, 8],
        "astar_3digits": [0.333, 0.181, 0.116, 0.082, 0.062, 0.049, 0.040, 0.033],
    },
}


def main():
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(FIXTURES, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
