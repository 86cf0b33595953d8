"""Convert Gaussian94-format basis files (Basis Set Exchange output, as shipped
in QUICK's ``basis/*.BAS``) into the compact JSON files this package loads.

Run once in a container that has the reference checkout (a provenance
recipe: nothing at run time reads the reference):

    python tests/golden/extract_bse.py /root/reference/basis

The JSON keeps the published exponents and *unnormalised* contraction
coefficients exactly as printed (Fortran ``D`` exponents converted to ``E``);
normalisation happens at load time (``quick_amd.basis``).  Shells are kept as
printed (S, SP, P, D, F) for the elements listed in ``ELEMENTS``.
"""
import json
import os
import sys

# basis keyword -> file name (QUICK's basis/basis_link table)
FILES = {
    "STO-3G": "STO-3G.BAS",
    "3-21G": "3-21G.BAS",
    "6-31G": "6-31G.BAS",
    "6-31G*": "6-31GD.BAS",
    "6-31G**": "6-31GDP.BAS",
    "cc-pVDZ": "CC-PVDZ.BAS",
    "def2-SVP": "DEF2-SVP.BAS",
    "cc-pVTZ": "CC-PVTZ.BAS",
    "def2-TZVP": "DEF2-TZVP.BAS",
}

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
            "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def parse_g94(path):
    out = {}
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    i = 0
    n = len(lines)
    while i < n:
        toks = lines[i].split()
        if len(toks) == 2 and toks[1] == "0" and toks[0].isalpha():
            elem = toks[0].capitalize()
            shells = []
            i += 1
            while i < n and not lines[i].startswith("****"):
                st = lines[i].split()
                if not st:
                    i += 1
                    continue
                kind = st[0].upper()
                nprim = int(st[1])
                exps, c1, c2 = [], [], []
                for k in range(nprim):
                    row = lines[i + 1 + k].split()
                    exps.append(_num(row[0]))
                    c1.append(_num(row[1]))
                    if kind == "SP":
                        c2.append(_num(row[2]))
                coefs = [c1, c2] if kind == "SP" else [c1]
                shells.append({"type": kind, "exps": exps, "coefs": coefs})
                i += 1 + nprim
            out[elem] = shells
        i += 1
    return out


def main(src_dir, dst_dir):
    for name, fname in FILES.items():
        data = parse_g94(os.path.join(src_dir, fname))
        keep = {e: data[e] for e in ELEMENTS if e in data}
        doc = {"name": name, "source": "Basis Set Exchange (Gaussian94 format)",
               "elements": keep}
        tag = name.replace("*", "s").replace("(", "").replace(")", "").replace(",", "")
        with open(os.path.join(dst_dir, tag + ".json"), "w") as fh:
            json.dump(doc, fh, indent=0, separators=(",", ":"))
        print(name, "->", tag + ".json", len(keep), "elements")


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basis"
    here = os.path.dirname(os.path.abspath(__file__))
    main(src, os.path.join(os.path.dirname(os.path.dirname(here)), "quick_amd", "basis_data"))
