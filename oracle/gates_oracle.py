"""CPU restatement of the reference's ancilla gates (christopherkang/2022-qpesim,
cpp/state_vector.hpp:51-113; Gate1D and its constructors, cpp/gate.hpp:7-39).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the product's
gate kernels.  The product never imports it.

The register is [2^m][D] (D amplitudes per block).  Qubit q is bit
t = m - q - 1 of the block index (big-endian, state_vector.hpp:56 and 84-85);
a gate mixes each block j0 with bit t clear and j1 = j0 | 2^t, amplitude by
amplitude: a0' = g00 a0 + g01 a1, a1' = g10 a0 + g11 a1.  A controlled gate
acts on the pairs whose control bit m - control - 1 is set."""
import numpy as np


def apply_gate(state, m, gate, qubit, control=None):
    """state: complex array of shape (2^m, D), modified in place."""
    g = np.asarray(gate, dtype=np.complex128).reshape(2, 2)
    t = m - qubit - 1
    c = None if control is None else m - control - 1
    for j0 in range(1 << m):
        if (j0 >> t) & 1:
            continue
        if c is not None and not (j0 >> c) & 1:
            continue
        j1 = j0 | (1 << t)
        a, b = state[j0].copy(), state[j1].copy()
        state[j0] = g[0, 0] * a + g[0, 1] * b
        state[j1] = g[1, 0] * a + g[1, 1] * b
    return state
