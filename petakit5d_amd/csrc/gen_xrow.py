#!/usr/bin/env python3
"""Writes pk_kern_xrow_*.hip: the (row length L, values per thread E) pairs
that get a compile-time row kernel, one per 2/3/5-smooth L in [32, 1024] that a rule below covers
(even rows of n1 = 2L), and one per odd 2/3/5-smooth n1 in [15, 1023] (odd rows, key -n1,
transformed at full length).

The first rule keeps every stage on home registers: E holds every prime of L
and P = L / E is a power of two dividing 256 (whole 256-thread blocks of
256 / P rows); the E with the lowest score 10 * stages + 4 * |E - 8| wins
(the host planner's rule).  For L <= 512 where that rule finds nothing, or only E > 16
(whose kernels need so many registers that one wave per SIMD fits), the
second rule takes any E with P <= 64 (floor(64 / P) rows per wavefront, the
64 mod P lanes left idle) or P = 128 / 256, and lets a stage take a radix
that does not divide E (such a stage runs from LDS); it scores
10 * stages + 10 * LDS-fed stages + 4 * |E - 8| + 100 x the idle-lane
fraction.  Run from this directory after changing a rule; the build compiles
whatever the files list."""
import os

E_LIST = [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32]
RADICES = [32, 30, 27, 25, 24, 20, 18, 16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2]
EXTRA = [(256, 4), (256, 16), (128, 4), (512, 16)]  # launch-shape sweep alternatives (PK_FFT_EX)


def smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def stages(L, E):
    n, rem = 0, L
    while rem > 1:
        r = next((r for r in RADICES if E % r == 0 and rem % r == 0), 0)
        if not r:
            return None
        rem //= r
        n += 1
    return n


def ct_stages(L, E):
    """Stages and LDS-fed stages of the compile-time rule (ct_radix_for)."""
    n, mm, rem = 0, 0, L
    while rem > 1:
        r = next((r for r in RADICES if E % r == 0 and rem % r == 0), 0)
        if not r:
            r = next(r for r in RADICES if rem % r == 0)
            mm += 1
        rem //= r
        n += 1
    return n, mm


def pick_registers(L):
    best = None
    for E in E_LIST:
        if L % E or any(L % p == 0 and E % p for p in (2, 3, 5)):
            continue
        P = L // E
        if P & (P - 1) or 256 % P:
            continue
        nst = stages(L, E)
        if nst is None:
            continue
        sc = nst * 10 + 4 * abs(E - 8)
        if best is None or sc < best[0]:
            best = (sc, E)
    return None if best is None else best[1]


def pick_mixed(L):
    best = None
    for E in E_LIST:
        if L % E:
            continue
        P = L // E
        if P > 64 and P not in (128, 256):
            continue
        nst, mm = ct_stages(L, E)
        sc = nst * 10 + mm * 10 + 4 * abs(E - 8) + (int(100 * (64 % P) / 64) if P <= 64 else 0)
        if best is None or sc < best[0]:
            best = (sc, E)
    return None if best is None else best[1]


def pick(L):
    E = pick_registers(L)
    if (E and E <= 16) or L > 512:
        return E
    return pick_mixed(L) or E


def main():
    pairs = []
    for L in range(32, 1025):
        if smooth(L):
            E = pick(L)
            if E:
                pairs.append((L, E))
    pairs += [p for p in EXTRA if p not in pairs]
    # odd rows (n1 = L odd, full-length complex transform): keyed -L
    for L in range(15, 1024, 2):
        if smooth(L):
            E = pick(L)
            if E:
                pairs.append((-L, E))
    here = os.path.dirname(os.path.abspath(__file__))
    for f in os.listdir(here):
        if f.startswith("pk_kern_xrow_") and f.endswith(".hip"):
            os.remove(os.path.join(here, f))
    per = 6
    groups = [pairs[i:i + per] for i in range(0, len(pairs), per)]
    names = []
    for gi, grp in enumerate(groups):
        tag = f"{gi:02d}"
        names.append(tag)
        def targs(L, E):
            return f"{L}, {E}" if L > 0 else f"{-L}, {E}, true"
        body = "\n".join(f"    if (L == {L} && E == {E}) {{\n        launch_modes_xrow<{targs(L, E)}>(grid, lds, s, a, tw);\n"
                         f"        return true;\n    }}" for L, E in grp)
        bodyb = "\n".join(f"    if (L == {L} && E == {E}) {{\n        launch_modes_xrowb<{targs(L, E)}>(grid, lds, s, mode, "
                          f"ch, starts, n);\n        return true;\n    }}" for L, E in grp)
        avail = " || ".join(f"(L == {L} && E == {E})" for L, E in grp)
        lst = ", ".join(f"{L}/{E}" for L, E in grp)
        with open(os.path.join(here, f"pk_kern_xrow_{tag}.hip"), "w") as fh:
            fh.write(f"""// pk_kern_xrow_{tag}.hip -- compile-time row-transform kernels, single-volume and batched
// launches, for (row length / values per thread) {lst}.  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {{
bool launch_xrow_{tag}(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {{
{body}
    return false;
}}
bool launch_xrowb_{tag}(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                      const int64_t* starts, int n) {{
{bodyb}
    return false;
}}
bool xrow_available_{tag}(int L, int E) {{ return {avail}; }}
}}  // namespace pk
""")
    decl = "\n".join(
        f"bool launch_xrow_{t}(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw);\n"
        f"bool launch_xrowb_{t}(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,\n"
        f"                      const int64_t* starts, int n);\n"
        f"bool xrow_available_{t}(int L, int E);" for t in names)
    one = " ||\n           ".join(f"launch_xrow_{t}(L, E, grid, lds, s, a, tw)" for t in names)
    oneb = " ||\n           ".join(f"launch_xrowb_{t}(L, E, grid, lds, s, mode, ch, starts, n)" for t in names)
    av = " ||\n           ".join(f"xrow_available_{t}(L, E)" for t in names)
    with open(os.path.join(here, "pk_kern_xrow_dispatch.hip"), "w") as fh:
        fh.write(f"""// pk_kern_xrow_dispatch.hip -- dispatch over the compile-time row kernels listed in
// pk_kern_xrow_NN.hip ({len(pairs)} (L, E) pairs).  Written by gen_xrow.py.
#include "pk_launch.hpp"

namespace pk {{
{decl}

bool launch_xrow(int L, int E, dim3 grid, size_t lds, hipStream_t s, const XArgs& a, const float2* tw) {{
    return {one};
}}
bool launch_xrow_batch(int L, int E, dim3 grid, size_t lds, hipStream_t s, int mode, const XChunk* ch,
                       const int64_t* starts, int n) {{
    return {oneb};
}}
bool xrow_available(int L, int E) {{
    return {av};
}}
}}  // namespace pk
""")
    print(f"{len(pairs)} pairs in {len(names)} files")


if __name__ == "__main__":
    main()
