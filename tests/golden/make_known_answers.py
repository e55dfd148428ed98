"""Write tests/golden/known_answers.json: perfect codes as known answers.

A perfect R-covering code C of X tiles X with radius-R balls, so
cnt(x) == 1 for every x, gain == 0 everywhere and loss(c) == |B| at every
codeword.  These facts come from coding theory, not from any code in this
repository or the reference, so they pin the oracle and the engine alike.

Codes (word index = sum_j digit_j q^j, digit j = coefficient of x^j):
* binary Hamming [7,4,3] and [15,11,3]: kernel of the parity-check matrix
  whose columns are 1..2^r-1 in binary;
* ternary Hamming [4,2,3]: parity-check columns (0,1),(1,0),(1,1),(1,2);
* ternary Golay [11,6,5]: cyclic, g(x) = x^5 + x^4 + 2x^3 + x^2 + 2;
* binary Golay [23,12,7]: cyclic, g(x) = x^11+x^10+x^6+x^5+x^4+x^2+1;
* Lee Z_5^2, R=1 (Golomb-Welch): {(i, 2i mod 5)};
* king torus 9x9, R=1: {(3i, 3j)};
* open 6x6 grid, R=1 (king domination): {(1,1),(1,4),(4,1),(4,4)}.
Every code is verified perfect by brute counting before it is written.
"""
import itertools
import json
import os

import numpy as np


def kernel_code(H, q):
    H = np.asarray(H, dtype=np.int64)
    n = H.shape[1]
    words = np.array(list(itertools.product(range(q), repeat=n)), dtype=np.int64)
    syn = (words @ H.T) % q
    return words[(syn == 0).all(axis=1)]


def cyclic_code(g, n, q):
    g = np.asarray(g, dtype=np.int64)        # coefficients, low degree first
    k = n - (len(g) - 1)
    out = []
    for m in itertools.product(range(q), repeat=k):
        c = np.zeros(n, dtype=np.int64)
        for i, mi in enumerate(m):
            if mi:
                c[i:i + len(g)] += mi * g
        out.append(c % q)
    return np.array(out)


def index(words, axes):
    pw = np.cumprod([1] + list(axes[:-1])).astype(np.int64)
    return (np.asarray(words, dtype=np.int64) @ pw).tolist()


def counts(problem, axes, R, code_digits):
    """cnt over X by brute counting (digit arrays, chunked)."""
    axes = np.asarray(axes)
    allw = np.array(list(itertools.product(*[range(a) for a in axes[::-1]])))[:, ::-1]
    code = np.asarray(code_digits)
    cnt = np.zeros(len(allw), dtype=np.int64)
    for lo in range(0, len(allw), 1 << 14):
        a = allw[lo:lo + (1 << 14), None, :]
        d = np.abs(a - code[None, :, :])
        if problem == "hamming":
            cov = (d != 0).sum(-1) <= R
        elif problem == "grid_linf":
            cov = d.max(-1) <= R
        else:
            d = np.minimum(d, axes - d)
            cov = (d.sum(-1) if problem == "lee" else d.max(-1)) <= R
        cnt[lo:lo + (1 << 14)] = cov.sum(1)
    return cnt


def main():
    cases = []

    def add(name, problem, axes, R, digits, ball):
        cnt = counts(problem, axes, R, digits)
        assert (cnt == 1).all(), name
        cases.append(dict(name=name, problem=problem, axes=list(axes), R=R,
                          ball=ball, code=index(digits, axes)))

    H7 = [[(c >> b) & 1 for c in range(1, 8)] for b in range(3)]
    add("binary Hamming [7,4]", "hamming", [2] * 7, 1, kernel_code(H7, 2), 8)
    H15 = [[(c >> b) & 1 for c in range(1, 16)] for b in range(4)]
    add("binary Hamming [15,11]", "hamming", [2] * 15, 1,
        kernel_code(H15, 2), 16)
    H3 = [[0, 1, 1, 1], [1, 0, 1, 2]]
    add("ternary Hamming [4,2]", "hamming", [3] * 4, 1, kernel_code(H3, 3), 9)
    add("ternary Golay [11,6]", "hamming", [3] * 11, 2,
        cyclic_code([2, 0, 1, 2, 1, 1], 11, 3), 243)
    lee = [(i, (2 * i) % 5) for i in range(5)]
    add("Lee Z5^2 Golomb-Welch", "lee", [5, 5], 1, lee, 5)
    king = [(3 * i, 3 * j) for i in range(3) for j in range(3)]
    add("king torus 9x9", "torus_linf", [9, 9], 1, king, 9)
    add("king grid 6x6", "grid_linf", [6, 6], 1,
        [(1, 1), (1, 4), (4, 1), (4, 4)], 9)
    golay23 = cyclic_code([1, 0, 1, 0, 1, 1, 1, 0, 0, 0, 1, 1], 23, 2)
    cases.append(dict(name="binary Golay [23,12]", problem="hamming",
                      axes=[2] * 23, R=3, ball=2048,
                      code=index(golay23, [2] * 23)))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "known_answers.json")
    with open(out, "w") as f:
        json.dump({"perfect_codes": cases}, f)
    print(out, [c["name"] for c in cases])


if __name__ == "__main__":
    main()
