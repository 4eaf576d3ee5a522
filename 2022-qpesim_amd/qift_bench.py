"""Time the QIFT kernels (qpe_qift) of libqpesim.so on one GPU.

For each dtype, ancilla count m and QIFT form (QPE_OPT_QIFT_FORM) it builds a
register of about --gb GB, runs the QIFT --reps times and prints the fastest
hipEvent time with the achieved HBM rate (2 * sizeof(amplitude) * 2^m * D bytes).
The register is zero-filled: the QIFT does the same work on any values.

    python 2022-qpesim_amd/qift_bench.py --m 4,5,6 --forms 1,2
"""
import argparse
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import qpesim  # noqa: E402

HBM_PEAK_TBS = 8.0


def block_shape(m, elem, gb):
    """(n, e) whose register of 2^m * C(n,e)^2 amplitudes is the largest not above gb"""
    best = None
    for n in range(4, 33):
        for e in range(1, n):
            C = math.comb(n, e)
            if C > 65535:
                continue
            size = (1 << m) * C * C * elem / 1e9
            if size <= gb and (best is None or size > best[0]):
                best = (size, n, e)
    return best[1], best[2]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--m", default="4,5,6,7,8,10")
    ap.add_argument("--forms", default="1", help="QIFT forms: 1 register DFTs, 2 matrix cores, 3 LDS radix-2")
    ap.add_argument("--dtypes", default="c64,c128")
    ap.add_argument("--gb", type=float, default=24.0)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    for name in a.dtypes.split(","):
        dtype, elem = (qpesim.COMPLEX64, 8) if name == "c64" else (qpesim.COMPLEX128, 16)
        for m in (int(x) for x in a.m.split(",")):
            n, e = block_shape(m, elem, a.gb)
            sv = qpesim.StateVector(m, n, e, dtype)
            D = sv.D
            for form in (int(x) for x in a.forms.split(",")):
                if form == qpesim.QIFT_MFMA and not 4 <= m <= 6:
                    continue
                sv.set_option(qpesim.OPT_QIFT_FORM, form)
                times = []
                for _ in range(a.reps + 1):
                    sv.qift()
                    times.append(sv.timings()["qift_ms"])
                t = min(times[1:])
                tbs = 2 * elem * (1 << m) * D / t / 1e9
                print(f"{name} form={form} m={m} D={D} qift_ms={t:.3f} TB/s={tbs:.2f} frac={tbs / HBM_PEAK_TBS:.3f}",
                      flush=True)
            sv.close()


if __name__ == "__main__":
    main()
