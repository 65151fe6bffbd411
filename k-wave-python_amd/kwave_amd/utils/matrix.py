"""Array padding (mirrors kwave/utils/matrix.py:94-156)."""
import numpy as np


def expand_matrix(matrix, exp_coeff, edge_val=None):
    """Pad ``matrix`` by ``exp_coeff`` cells per side, repeating edges or with ``edge_val``."""
    matrix = np.squeeze(np.asarray(matrix))
    opts = {"mode": "edge"} if edge_val is None else {"mode": "constant", "constant_values": edge_val}
    c = np.array(exp_coeff).astype(int).squeeze()
    n = c.size
    if n == 0:
        raise ValueError("exp_coeff must not be empty")
    nd = matrix.ndim
    if n == 1 or nd == 1:
        pw = c
    elif nd == 2 and n == 2:
        pw = [(c[0],), (c[1],)]
    elif nd == 2 and n == 4:
        pw = [(c[0], c[1]), (c[2], c[3])]
    elif nd == 3 and n == 3:
        pw = [(c[0], c[0]), (c[1], c[1]), (c[2], c[2])]
    elif nd == 3 and n == 6:
        pw = [(c[0], c[1]), (c[2], c[3]), (c[4], c[5])]
    else:
        raise ValueError(f"exp_coeff of {n} elements does not fit a {nd}-D matrix")
    return np.pad(matrix, pad_width=pw, **opts)


def num_dim2(x):
    """Number of non-singleton dimensions (at least 2 for matrices)."""
    return int(np.ndim(np.squeeze(x)))
