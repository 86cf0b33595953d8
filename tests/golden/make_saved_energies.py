"""Collect QUICK's own RHF and UHF regression fixtures into small JSON files.

Reads the reference's ``test/ene_*_rhf_*.in`` inputs (geometry, basis,
charge) and the matching ``test/saved/*.out`` outputs (TOTAL ENERGY,
ELECTRONIC ENERGY, CORE_CORE REPULSION, the SCF cycle count, and the
TOTAL 2e TIME of the Fock builds with the GPU it ran on) and writes
``saved_rhf_energies.json`` beside this script.  The ``*_uhf_*.in`` inputs
(with their multiplicity) go to ``saved_uhf_energies.json``; for a geometry
optimisation the record is the first SCF, at the input geometry.  Only data
is kept: inputs and expected outputs.

    python tests/golden/make_saved_energies.py /root/reference/test
"""
import glob
import json
import os
import re
import sys

BASES = {"STO-3G": "STO-3G", "3-21G": "3-21G", "6-31G": "6-31G", "6-31G*": "6-31G*",
         "6-31G**": "6-31G**", "cc-pvdz": "cc-pVDZ", "6-31g**": "6-31G**"}
KEEP = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
        "P", "S", "Cl", "Ar"}


def parse_in(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    kw = lines[0]
    m = re.search(r"BASIS=(\S+)", kw)
    basis = m.group(1)
    m = re.search(r"CHARGE=([+-]?\d+)", kw)
    charge = int(m.group(1)) if m else 0
    m = re.search(r"MULT=(\d+)", kw, re.I)
    mult = int(m.group(1)) if m else 1
    m = re.search(r"cutoff=(\S+)", kw, re.I)
    cutoff = float(m.group(1)) if m else 1e-7
    syms, xyz = [], []
    for ln in lines[2:]:
        t = ln.split()
        if len(t) != 4:
            if syms:
                break
            continue
        syms.append(t[0].capitalize())
        xyz.append([float(v) for v in t[1:]])
    return basis, charge, cutoff, syms, xyz, kw, mult


def parse_out(path):
    out = {}
    with open(path) as fh:
        txt = fh.read()
    for key, tag in [("E_total", "TOTAL ENERGY"), ("E_elec", "ELECTRONIC ENERGY"),
                     ("E_nuc", "CORE_CORE REPULSION")]:
        m = re.search(tag + r"\s*=\s*(-?\d+\.\d+)", txt)
        if m:
            out[key] = float(m.group(1))
    m = re.search(r"TOTAL 2e TIME\s*=\s*(\d+\.\d+)", txt)
    if m:
        out["t2e_total_s"] = float(m.group(1))      # host wall time of all Fock builds of the run
    m = re.search(r"(?:CUDA|HIP) DEVICE NAME\s*:\s*(.+?)\s*$", txt, re.M)
    if m:
        out["device"] = m.group(1)
    m = re.search(r"REACH CONVERGENCE AFTER\s+(\d+)", txt)
    if m:
        out["cycles"] = int(m.group(1))
    m = re.findall(r"BASIS FUNCTIONS =\s+(\d+)", txt)
    if m:
        out["nbasis"] = int(m[-1])   # the last one is the molecule (earlier ones: SAD atoms)
    return out


def collect(test_dir, pattern):
    recs = []
    for inp in sorted(glob.glob(os.path.join(test_dir, pattern))):
        name = os.path.basename(inp)[:-3]
        outp = os.path.join(test_dir, "saved", name + ".out")
        if not os.path.exists(outp):
            continue
        basis, charge, cutoff, syms, xyz, kw, mult = parse_in(inp)
        if basis not in BASES or not set(syms) <= KEEP:
            continue
        if "DFT" in kw.upper() or "B3LYP" in kw.upper():
            continue
        res = parse_out(outp)
        if "E_total" not in res:
            continue
        rec = dict(name=name, basis=BASES[basis], charge=charge, cutoff=cutoff,
                   symbols=syms, xyz_angstrom=xyz, **res)
        if mult != 1:
            rec["mult"] = mult
        recs.append(rec)
    return recs


def main(test_dir):
    for pattern, fname in [("ene_*_rhf_*.in", "saved_rhf_energies.json"), ("*_uhf_*.in", "saved_uhf_energies.json")]:
        recs = collect(test_dir, pattern)
        dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), fname)
        with open(dst, "w") as fh:
            json.dump({"source": "QUICK test/saved regression outputs (printed to 1e-9 Eh)",
                       "records": recs}, fh, indent=1)
        print(len(recs), "records ->", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test")
