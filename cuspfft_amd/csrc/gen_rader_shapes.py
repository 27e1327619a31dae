"""Generate rader_shapes.h: the k_rowrader shapes for every prime p whose p - 1 is
7-smooth, for odd prime widths C = p (one convolution per row, NS = 1) with
1024 <= p <= 16384 and for widths C = 2p (two convolutions per row, NS = 2) with
1024 <= 2p <= 16384.

Each shape is X(M1, NT, R1, R2, R3, NS) with L = p - 1 = M1 * R1 * R2 * R3:
  * M1 (the register codelet of phases A and C) is a product of whole prime-power
    blocks of L, so gcd(M1, M2) = 1 (the kernels' prime-factor index maps need it),
    and 3 <= M1 <= 32;
  * M2 = R1 R2 R3 (the LDS Stockham stages, radix 1 = no stage) with NS * M2 <= 1024
    (one column n2 of one convolution per thread);
  * NT = NS * M2 rounded up to whole waves (at least 128 threads);
  * the row buffer NS * M1 * M2 entries (plus tables) within 160 KiB of LDS;
  * registers: 5 * M1 + 40 (phases A and C; measured 150-184 VGPRs at M1 = 25-32)
    and, per stage, (butterflies per thread + 1) * 2 * R VGPRs, within the budget of
    NT threads' occupancy (4, 3, 2 waves per SIMD at 1024, 768, 512 threads).
Among the valid shapes the one with the smallest cost estimate is taken: the codelet
operations of both directions plus 16 per entry and LDS pass (every stage is one
LDS round trip of the row buffer), which favours two stages over three, weighted by
1 + 128 / (NS M2): the codelet phases run on NS M2 threads only.

Usage: python3 gen_rader_shapes.py   (rewrites rader_shapes.h next to this file)
"""
import itertools
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LDS_LIMIT = 160 * 1024
MAX_RADIX = 16
LISTS = 4  # generated lists (one compilation unit each)
# widths whose chosen shape measured slower than the Bluestein plan of the same matrix
# (whole transform, square matrix with 5 nonzeros per row, one MI355X): left out so that
# AUTO keeps Bluestein there
EXCLUDE = {3001: "(3, 1024, 10, 10, 10, 1): 0.132 ms against Bluestein's 0.121 ms"}


def is_prime(n):
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def factor(n):
    f = {}
    d = 2
    while d * d <= n:
        while n % d == 0:
            f[d] = f.get(d, 0) + 1
            n //= d
        d += 1
    if n > 1:
        f[n] = f.get(n, 0) + 1
    return f


def codelet_ops(n):
    """Rough real-operation count of the register codelet of length n (7-smooth)."""
    ops = 0.0
    for q, e in factor(n).items():
        per = {2: 2.0, 3: 4.0, 5: 7.0, 7: 10.0}.get(q, 2.0 * q)
        ops += e * per
    return n * ops + 6.0 * n  # butterflies + twiddle products between the codelet's stages


def radix_splits(m2):
    """Ordered factorisations of m2 into 1-3 radices in [2, MAX_RADIX] (non-increasing)."""
    out = []
    if m2 <= MAX_RADIX:
        out.append((m2, 1, 1))
    for r1 in range(2, MAX_RADIX + 1):
        if m2 % r1:
            continue
        rest = m2 // r1
        if 2 <= rest <= MAX_RADIX and rest <= r1:
            out.append((r1, rest, 1))
        for r2 in range(2, min(r1, MAX_RADIX) + 1):
            if rest % r2:
                continue
            r3 = rest // r2
            if 2 <= r3 <= r2:
                out.append((r1, r2, r3))
    return out


def vgpr_cap(nt):
    waves = -(-nt // 64)
    per_simd = -(-waves // 4)
    return {1: 256, 2: 256, 3: 168, 4: 128}.get(per_simd, 128)


def lds_bytes(ns, m1, m2, radices):
    # W_M2, the r-major tables (at most 2 * M2 entries), the row buffer, partial sums
    buf = max(ns * m1 * m2, ns * (m1 * m2 + 1))
    return (m2 + 2 * m2 + buf + 16 * ns + 64) * 8


def shape_for(p, ns):
    L = p - 1
    blocks = [q ** e for q, e in factor(L).items()]
    best = None
    for k in range(1, len(blocks) + 1):
        for sub in itertools.combinations(blocks, k):
            m1 = math.prod(sub)
            m2 = L // m1
            if not (3 <= m1 <= 32) or m2 < 2 or ns * m2 > 1024:
                continue
            nt = max(128, -(-ns * m2 // 64) * 64)
            cap = vgpr_cap(nt)
            if 5 * m1 + 40 > cap:
                continue
            for rad in radix_splits(m2):
                stages = [r for r in rad if r > 1]
                ok = True
                for r in stages:
                    mb = -(-(ns * m1 * (m2 // r)) // nt)
                    if (mb + 1) * 2 * r + 24 > cap:
                        ok = False
                if not ok or lds_bytes(ns, m1, m2, rad) > LDS_LIMIT:
                    continue
                cost = ns * (2 * m2 * codelet_ops(m1) + 2 * m1 * sum((m2 // r) * codelet_ops(r) for r in stages)
                             + 16.0 * L * (2 * len(stages) + 1))
                # fewer live threads (NS M2) leave the codelet phases serial: weigh them
                key = (cost * (1.0 + 128.0 / (ns * m2)), -m1)
                if best is None or key < best[0]:
                    best = (key, (m1, nt, rad[0], rad[1], rad[2], ns))
    return best[1] if best else None


def shapes():
    out = []
    for p in range(3, 16385):
        if not is_prime(p) or max(factor(p - 1)) > 7:
            continue
        if 1024 <= p and p not in EXCLUDE:
            s = shape_for(p, 1)
            if s:
                out.append((p, s))
        if 1024 <= 2 * p <= 16384 and 2 * p not in EXCLUDE:
            s = shape_for(p, 2)
            if s:
                out.append((2 * p, s))
    return out


def render():
    """The text of rader_shapes.h."""
    sh = shapes()
    lists = [[] for _ in range(LISTS)]
    # spread by width so each list holds small and large rows alike
    for i, (w, s) in enumerate(sorted(sh)):
        lists[i % LISTS].append((w, s))
    lines = [
        "// rader_shapes.h -- generated by gen_rader_shapes.py; do not edit.",
        "// k_rowrader shapes X(M1, NT, R1, R2, R3, NS) for the primes p with 7-smooth p - 1:",
        "// odd prime widths C = p (NS = 1, 1024 <= p <= 16384) and widths C = 2p (NS = 2,",
        "// 1024 <= 2p <= 16384); p - 1 = M1 R1 R2 R3.  One list per compilation unit",
        "// (rowpass_rader_gen_{a,b,c,d}.hip).",
        "#pragma once",
        "",
    ]
    for li, name in enumerate("ABCD"):
        lines.append(f"#define SPFFT_RADER_GEN_LIST_{name}(X) \\")
        body = [f"    X({m1}, {nt}, {r1}, {r2}, {r3}, {ns}) /* C = {w} */" for w, (m1, nt, r1, r2, r3, ns) in lists[li]]
        lines.append(" \\\n".join(body))
        lines.append("")
    return "\n".join(lines)


def main():
    text = render()
    with open(os.path.join(HERE, "rader_shapes.h"), "w") as f:
        f.write(text)
    sh = shapes()
    print(f"{len(sh)} shapes ({sum(1 for w, s in sh if s[5] == 1)} odd prime widths, "
          f"{sum(1 for w, s in sh if s[5] == 2)} widths 2p)")


if __name__ == "__main__":
    main()
