"""Write gri30_ndodecane.mech: the compiled GRI-3.0 and n-dodecane mechanisms
side by side in one phase (53 + 100 = 153 species, 325 + 553 reactions), a
synthetic mechanism with more species than the dense integrators take.  The
second mechanism's species indices are offset by the first's species count and
the element compositions remapped to the union of the element lists; no
species name is shared (GRI-3.0's names are upper case, n-dodecane's lower
case).  Reactions keep their tables unchanged.

Usage: python tests/golden/make_merged_mech.py  (reads cantera_amd/mechanisms)
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MECH = os.path.join(HERE, "..", "..", "cantera_amd", "mechanisms")


def read(path):
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f]
    head = {}
    i = 0
    while not lines[i].startswith("species"):
        k, _, v = lines[i].partition(" ")
        head[k] = v
        i += 1
    et = head["elements"].split()
    elements = [(et[1 + 2 * e], et[2 + 2 * e]) for e in range(int(et[0]))]
    K = int(lines[i].split()[1])
    species = [lines[i + 1 + k].split() for k in range(K)]
    i += 1 + K
    assert lines[i].startswith("reactions")
    reactions = []          # (R tokens, following lines up to and including E)
    j = i + 1
    while lines[j] != "end":
        assert lines[j].startswith("R ")
        r = lines[j].split()
        extra = []
        j += 1
        while not lines[j].startswith("E "):
            extra.append(lines[j])
            j += 1
        extra.append(lines[j])
        j += 1
        reactions.append((r, extra))
    return elements, species, reactions


def offset_reaction(r, koff):
    """Species indices of an R line (reactants k s o, products k s, third-body
    k e) shifted by koff."""
    t = list(r)
    p = 15                   # R idx rtype rev dn A b EaR A0 b0 EaR0 troe[4]
    nr = int(t[p]); p += 1
    for _ in range(nr):
        t[p] = str(int(t[p]) + koff); p += 3
    npr = int(t[p]); p += 1
    for _ in range(npr):
        t[p] = str(int(t[p]) + koff); p += 2
    p += 1                   # tb_default
    ntb = int(t[p]); p += 1
    for _ in range(ntb):
        t[p] = str(int(t[p]) + koff); p += 2
    assert p == len(t), (p, len(t))
    return t


def main():
    ea, sa, ra = read(os.path.join(MECH, "gri30.mech"))
    eb, sb, rb = read(os.path.join(MECH, "nDodecane.mech"))
    names = [e for e, _ in ea] + [e for e, _ in eb if e not in [x for x, _ in ea]]
    aw = dict(ea + eb)
    names_a = [s[2] for s in sa]
    assert not set(names_a) & {s[2] for s in sb}
    out = ["CGMECH 1", "phase gri30_ndodecane", "source gri30.mech + nDodecane.mech (merged)",
           f"elements {len(names)} " + " ".join(f"{e} {aw[e]}" for e in names),
           f"species {len(sa) + len(sb)}"]
    for block, els in ((sa, ea), (sb, eb)):
        for s in block:
            comp = dict(zip([e for e, _ in els], s[-len(els):]))
            idx = len(out) - 5
            out.append(" ".join(["S", str(idx)] + s[2:-len(els)] + [comp.get(e, "0.0") for e in names]))
    out.append(f"reactions {len(ra) + len(rb)}")
    for koff, block in ((0, ra), (len(sa), rb)):
        for r, extra in block:
            t = offset_reaction(r, koff)
            t[1] = str(len([l for l in out if l.startswith("R ")]))
            out.append(" ".join(t))
            out.extend(extra)
    out.append("end")
    with open(os.path.join(HERE, "gri30_ndodecane.mech"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
