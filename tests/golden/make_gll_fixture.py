"""Write tests/golden/gll_tables.json: closed-form GLL / D / VCJH / filter
tables for p = 1..8 computed with mpmath at 50 digits.

These pin the setup tables of both the oracle and the product independently
of either implementation:
  * GLL nodes = {-1, roots of P_p', +1}; weights 2 / (p (p+1) P_p(x)^2);
  * D_ij = l_j'(x_i) for the Lagrange basis on the nodes;
  * VCJH: g_R' = (P_p' + (eta P_{p-1}' + P_{p+1}')/(1+eta)) / 2,
          g_L' = (-1)^p (P_p' - (eta P_{p-1}' + P_{p+1}')/(1+eta)) / 2,
    eta = c (2p+1) (a_p p!)^2 / 2, a_p = (2p)! / (2^p (p!)^2)
    (Vincent, Castonguay & Jameson 2011), for c in {0, 0.01};
  * filter F = V diag(sigma) V^-1 (modal exponential filter).
Run:  python tests/golden/make_gll_fixture.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gll_tables.json")


def P(n, x):
    return mp.legendre(n, x)


def dP(n, x):
    # P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1); P_n'(+-1) = (+-1)^(n-1) n (n+1) / 2
    if n == 0:
        return mp.mpf(0)
    if abs(abs(x) - 1) < mp.mpf(10) ** -40:
        return mp.sign(x) ** (n - 1) * n * (n + 1) / 2
    return n * (x * P(n, x) - P(n - 1, x)) / (x * x - 1)


def nodes(p):
    if p == 1:
        return [mp.mpf(-1), mp.mpf(1)]
    # interior: roots of P_p', each isolated by a sign change on a fine grid and
    # refined by bisection-type root finding at 50 digits
    grid = [mp.mpf(-1) + 2 * mp.mpf(k) / 3999 for k in range(1, 3999)]  # odd count: 0 is not a grid point
    xs = [mp.mpf(-1)]
    for a, b in zip(grid[:-1], grid[1:]):
        if mp.sign(dP(p, a)) != mp.sign(dP(p, b)):
            xs.append(mp.findroot(lambda t: dP(p, t), (a, b), solver="anderson"))
    xs.append(mp.mpf(1))
    assert len(xs) == p + 1, (p, len(xs))
    return sorted(xs)


def lagrange_deriv(xs):
    # D_ij = l_j'(x_i): (b_j / b_i) / (x_i - x_j) off the diagonal with b_j = 1/prod_{m!=j}(x_j - x_m),
    # D_ii = sum_{m != i} 1 / (x_i - x_m)  (exact in mpmath arithmetic)
    n = len(xs)
    b = []
    for j in range(n):
        r = mp.mpf(1)
        for m in range(n):
            if m != j:
                r *= xs[j] - xs[m]
        b.append(1 / r)
    D = [[mp.mpf(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i][j] = (b[j] / b[i]) / (xs[i] - xs[j])
            else:
                D[i][i] = sum(1 / (xs[i] - xs[m]) for m in range(n) if m != i)
    return D


def vcjh(p, c, xs):
    ap = mp.factorial(2 * p) / (2 ** p * mp.factorial(p) ** 2)
    eta = c * (2 * p + 1) * (ap * mp.factorial(p)) ** 2 / 2
    dR, dL = [], []
    for x in xs:
        corr = (eta * dP(p - 1, x) + dP(p + 1, x)) / (1 + eta)
        dR.append((dP(p, x) + corr) / 2)
        dL.append((-1) ** p * (dP(p, x) - corr) / 2)
    return dL, dR


def filt(p, xs, n_c, s, alpha):
    n = p + 1
    V = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            V[i, j] = P(j, xs[i])
    S = mp.matrix(n, n)
    for i in range(n):
        S[i, i] = 1 if i < n_c else mp.exp(-alpha * (mp.mpf(i - n_c) / ((n - 1) - n_c)) ** s)
    F = V * S * V ** -1
    return [[F[i, j] for j in range(n)] for i in range(n)]


def f(x):
    return float(x)


def main():
    out = {}
    for p in range(1, 9):
        xs = nodes(p)
        w = [2 / (p * (p + 1) * P(p, x) ** 2) for x in xs]
        D = lagrange_deriv(xs)
        dL0, dR0 = vcjh(p, 0, xs)
        dL1, dR1 = vcjh(p, mp.mpf("0.01"), xs)
        F = filt(p, xs, 0, 16, 36) if p >= 1 else None
        F2 = filt(p, xs, 1, 32, 36) if p >= 2 else None
        out[str(p)] = dict(
            csi=[f(x) for x in xs], w=[f(x) for x in w],
            D=[[f(v) for v in row] for row in D],
            dL_c0=[f(v) for v in dL0], dR_c0=[f(v) for v in dR0],
            dL_c001=[f(v) for v in dL1], dR_c001=[f(v) for v in dR1],
            filt_nc0_s16_a36=[[f(v) for v in row] for row in F],
            filt_nc1_s32_a36=[[f(v) for v in row] for row in F2] if F2 else None,
        )
        print("p", p, "done", flush=True)
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
