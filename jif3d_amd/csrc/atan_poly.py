"""Coefficients of the FP64 arctangent used by the corner terms (prism_device.hpp, atan_f64).

atan(v) = v + v^3 g(v^2) on |v| <= tan(pi/8); g is the degree-10 polynomial in x = v^2 that
minimises the largest relative error of atan (Remez exchange, 60 significant digits), rounded to
doubles.  Run `python jif3d_amd/csrc/atan_poly.py` to print them and the error they leave.
"""
import mpmath as mp

mp.mp.dps = 60
DEG = 10
T2 = mp.tan(mp.pi / 8) ** 2


def g(x):
    if x == 0:
        return mp.mpf(-1) / 3
    r = mp.sqrt(x)
    return (mp.atan(r) - r) / (x * r)


def weight(x):
    # relative error of atan(v) = v (1 + x g(x)): the error of g is scaled by x / (1 + x g)
    r = mp.sqrt(x)
    return x * r / mp.atan(r) if x > 0 else mp.mpf(0)


def remez():
    n = DEG + 1
    # initial reference: Chebyshev extrema on [0, T2]
    pts = [T2 * (1 - mp.cos(mp.pi * i / (n))) / 2 for i in range(n + 1)]
    pts[0] = T2 * mp.mpf("1e-6")
    for _ in range(40):
        # solve sum c_k x^k + (-1)^i E / w(x) = g(x) at the n + 1 points
        A = mp.matrix(n + 1, n + 1)
        b = mp.matrix(n + 1, 1)
        for i, x in enumerate(pts):
            for k in range(n):
                A[i, k] = x ** k
            A[i, n] = (-1) ** i / weight(x)
            b[i] = g(x)
        sol = mp.lu_solve(A, b)
        c = [sol[k] for k in range(n)]
        # new extrema of the weighted error
        def err(x):
            return (mp.polyval(c[::-1], x) - g(x)) * weight(x)
        grid = [T2 * mp.mpf(i) / 4000 for i in range(1, 4001)]
        vals = [err(x) for x in grid]
        ext = [grid[0]]
        for i in range(1, len(grid) - 1):
            if (vals[i] - vals[i - 1]) * (vals[i + 1] - vals[i]) <= 0:
                ext.append(grid[i])
        ext.append(grid[-1])
        ext = sorted(ext, key=lambda x: -abs(err(x)))[:n + 1]
        pts = sorted(ext)
    return c, max(abs(v) for v in vals)


if __name__ == "__main__":
    c, e = remez()
    print(f"# max relative error of atan from the polynomial (exact arithmetic): {mp.nstr(e, 3)}")
    for k, ck in enumerate(c):
        print(f"c{k} = {float(ck)!r}")
