"""Small-case arithmetic in Python integers, written independently of the
oracle's formulation, used only to pin the oracle on tiny inputs."""


def negacyclic_mul(a, b, q):
    d = len(a)
    out = [0] * d
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            if k < d:
                out[k] += x * y
            else:
                out[k - d] -= x * y
    return [v % q for v in out]


def eval_poly(a, x, q):
    acc = 0
    for c in reversed(a):
        acc = (acc * x + c) % q
    return acc


def eq_weights(cs, q):
    """weights[y] = prod_b (c_b if bit_b(y) else 1 - c_b), first c = MSB."""
    n = len(cs)
    out = []
    for y in range(1 << n):
        w = 1
        for b, c in enumerate(cs):
            bit = (y >> (n - 1 - b)) & 1
            w = w * (c if bit else (1 - c)) % q
        out.append(w)
    return out


class Fq2:
    """F_q[Y]/(Y^2 - nr) with Python ints."""

    def __init__(self, q, nr):
        self.q, self.nr = q, nr

    def mul(self, x, y):
        q, nr = self.q, self.nr
        return ((x[0] * y[0] + nr * x[1] * y[1]) % q, (x[0] * y[1] + x[1] * y[0]) % q)

    def add(self, x, y):
        return ((x[0] + y[0]) % self.q, (x[1] + y[1]) % self.q)

    def sub(self, x, y):
        return ((x[0] - y[0]) % self.q, (x[1] - y[1]) % self.q)


def mle_eval(w, cs, q, f2=None):
    """Multilinear extension of w at cs (first challenge = MSB), by repeated
    halving written as an explicit interpolation w0*(1-c) + w1*c."""
    if f2 is None:
        cur = [x % q for x in w]
        for c in cs:
            h = len(cur) // 2
            cur = [(cur[x] * (1 - c) + cur[x + h] * c) % q for x in range(h)]
        return cur[0]
    cur = [(x % q, 0) for x in w]
    one = (1, 0)
    for c in cs:
        h = len(cur) // 2
        omc = f2.sub(one, c)
        cur = [f2.add(f2.mul(cur[x], omc), f2.mul(cur[x + h], c)) for x in range(h)]
    return cur[0]


def round_poly_brute(w, cs_prefix, q, points=(0, 1, 2)):
    """g(X) = sum_{x'} W(cs_prefix, X, x')^2 at X in points, by evaluating the
    multilinear extension on the boolean suffix hypercube."""
    n = (len(w) - 1).bit_length()
    t = len(cs_prefix)
    rest = n - t - 1
    vals = []
    for X in points:
        s = 0
        for suffix in range(1 << rest):
            bits = [(suffix >> (rest - 1 - b)) & 1 for b in range(rest)]
            s += mle_eval(w, list(cs_prefix) + [X] + bits, q) ** 2
        vals.append(s % q)
    return vals
