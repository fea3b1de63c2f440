"""A first-GN-iteration system of the oracle as a 6x6 block-CSR (test helper)."""
import numpy as np


def block_csr_of(res, N):
    """Row i = [the diagonal block, then the off-diagonal blocks of its
    symmetric keys (i * 4096 + j) in key order], row-major 6x6 blocks; also
    the FP64 inverse diagonal blocks (N x 36) and the right-hand side."""
    keys = res["sym_keys"].astype(np.int64)
    off = res["offdiag"].reshape(-1, 36)
    D = res["diag"].reshape(N, 36)
    rows = keys // 4096
    row_ptr = np.zeros(N + 1, np.uint32)
    col, blocks = [], []
    for i in range(N):
        sel = np.nonzero(rows == i)[0]
        col += [i] + list(keys[sel] % 4096)
        blocks += [D[i]] + [off[k] for k in sel]
        row_ptr[i + 1] = len(col)
    return (row_ptr, np.array(col, np.uint32), np.array(blocks, np.float64), res["diag_inv"].reshape(N, 36),
            res["rhs"].astype(np.float64))


def dense_of(row_ptr, col, blocks):
    N = len(row_ptr) - 1
    A = np.zeros((6 * N, 6 * N))
    for i in range(N):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            A[6 * i:6 * i + 6, 6 * col[k]:6 * col[k] + 6] = blocks[k].reshape(6, 6)
    return A
