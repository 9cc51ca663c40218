"""Polynomial coefficients of the FP64 corner terms' elementary functions (prism_device.hpp).

atan (atan2_f64): atan(v) = v + v^3 g(v^2) on |v| <= tan(pi/8); g is the degree-10 polynomial in
x = v^2 that minimises the largest relative error of atan.

log (log_pos_f64): log m = 2 atanh(f) = 2 f + 2 f x q(x), x = f^2, |f| <= (sqrt2 - 1)/(sqrt2 + 1);
q is the degree-7 polynomial that minimises its own largest relative error (q(x) ~ 1/3 + x/5 +
...; its error enters the logarithm scaled by about x q <= 0.01).

Both by Remez exchange at 60 significant digits, rounded to doubles.  Run
`python jif3d_amd/csrc/corner_poly.py` to print them and the errors they leave.
"""
import mpmath as mp

mp.mp.dps = 60
ATAN_DEG = 10
LOG_DEG = 7


def remez(fun, weight, top, deg, start=mp.mpf(0), iters=40):
    """Coefficients c_0..c_deg minimising max |(sum c_k x^k - fun(x)) weight(x)| on [start, top]."""
    n = deg + 1
    pts = [start + (top - start) * (1 - mp.cos(mp.pi * i / n)) / 2 for i in range(n + 1)]
    if weight(pts[0]) == 0:
        pts[0] = top * mp.mpf("1e-6")
    vals = []
    for _ in range(iters):
        A = mp.matrix(n + 1, n + 1)
        b = mp.matrix(n + 1, 1)
        for i, x in enumerate(pts):
            for k in range(n):
                A[i, k] = x ** k
            A[i, n] = (-1) ** i / weight(x)
            b[i] = fun(x)
        sol = mp.lu_solve(A, b)
        c = [sol[k] for k in range(n)]

        def err(x):
            return (mp.polyval(c[::-1], x) - fun(x)) * weight(x)
        grid = [start + (top - start) * mp.mpf(i) / 4000 for i in range(1, 4001)]
        vals = [err(x) for x in grid]
        ext = [grid[0]]
        for i in range(1, len(grid) - 1):
            if (vals[i] - vals[i - 1]) * (vals[i + 1] - vals[i]) <= 0:
                ext.append(grid[i])
        ext.append(grid[-1])
        pts = sorted(sorted(ext, key=lambda x: -abs(err(x)))[:n + 1])
    return c, max(abs(v) for v in vals)


def atan_g(x):
    if x == 0:
        return mp.mpf(-1) / 3
    r = mp.sqrt(x)
    return (mp.atan(r) - r) / (x * r)


def atan_weight(x):
    # relative error of atan(v) = v (1 + x g(x)): the error of g is scaled by x v / atan(v)
    r = mp.sqrt(x)
    return x * r / mp.atan(r) if x > 0 else mp.mpf(0)


def log_q(x):
    if x == 0:
        return mp.mpf(1) / 3
    r = mp.sqrt(x)
    return (mp.atanh(r) / r - 1) / x


def main():
    c, e = remez(atan_g, atan_weight, mp.tan(mp.pi / 8) ** 2, ATAN_DEG)
    print(f"# atan: max relative error of atan from g (exact arithmetic) {mp.nstr(e, 3)}")
    for k, ck in enumerate(c):
        print(f"atan c{k} = {float(ck)!r}")
    top = ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2
    c, e = remez(log_q, lambda x: 1 / log_q(x), top, LOG_DEG, iters=30)
    print(f"# log: max relative error of q {mp.nstr(e, 3)} (in the logarithm <= {mp.nstr(e * top, 3)})")
    for k, ck in enumerate(c):
        print(f"log c{k} = {float(ck)!r}")


if __name__ == "__main__":
    main()
