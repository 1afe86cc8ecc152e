"""Derive 6-31g.gbs from 6-31gs.gbs (6-31G* = 6-31G + one D shell on heavy atoms).

The s/sp contractions of 6-31G and 6-31G* are identical, so 6-31G is obtained by
dropping every D shell block.  Run from this directory:  python make_631g.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def derive(src=os.path.join(HERE, "6-31gs.gbs"), dst=os.path.join(HERE, "6-31g.gbs")):
    out = []
    with open(src) as fh:
        lines = fh.read().split("\n")
    skip = 0
    for line in lines:
        if skip:
            skip -= 1
            continue
        f = line.split()
        if len(f) == 3 and f[0] == "D" and f[1].isdigit():
            skip = int(f[1])
            continue
        out.append(line.replace("6-31G*", "6-31G (derived: 6-31G* without D shells)", 1)
                   if line.startswith("!   Basis set:") else line)
    with open(dst, "w") as fh:
        fh.write("\n".join(out))
    return dst


if __name__ == "__main__":
    print(derive())
