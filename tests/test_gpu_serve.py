"""Continuous-batching SCF server: every request's energy is the same when
served alone and in a mix of concurrently running jobs (mixed precision,
tolerance 1e-8 Eh), and within 1e-6 Eh of the FP64 SCF of the oracle."""
import numpy as np
import pytest

from tests.helpers import announce

pytestmark = pytest.mark.gpu


def test_served_alone_equals_served_in_the_mix():
    announce("serving: alone vs mix")
    import torch
    from quick_amd import molecules as M
    from quick_amd.serve import FockServer, Request
    mols = [(M.water(), "6-31G*"), (M.alkane(2), "6-31G"), (M.water_cluster(3), "def2-SVP"),
            (M.alkane(1), "6-31G*"), (M.water_cluster(2), "6-31G")]
    dev = torch.device("cuda:0")
    # in the mix: staggered arrivals, a budget that admits ~2 jobs at a time
    reqs = [Request(i, m, b, arrival=0.02 * i) for i, (m, b) in enumerate(mols)]
    from quick_amd.serve import footprint
    budget = 2.2 * max(footprint(r) for r in reqs)
    srv = FockServer(dev, hbm_budget=budget, precision="mixed")
    mix, stats = srv.run(reqs)
    assert len(mix) == len(reqs) and all(r.converged for r in mix)
    mix = {r.rid: r for r in mix}
    for i, (m, b) in enumerate(mols):
        alone, _ = FockServer(dev, precision="mixed").run([Request(i, m, b, 0.0)])
        dE = abs(alone[0].energy - mix[i].energy)
        print(f"  req {i} {b}: E {mix[i].energy:.10f}  |alone - mix| {dE:.1e}", flush=True)
        assert dE <= 1e-8


def test_served_energy_matches_oracle_scf():
    announce("serving: water 6-31G* vs oracle SCF")
    import torch
    from oracle import oracle as O
    from quick_amd import molecules as M
    from quick_amd.serve import FockServer, Request
    mol = M.water()
    res, _ = FockServer(torch.device("cuda:0"), precision="mixed").run([Request(0, mol, "6-31G*", 0.0)])
    qb = O.QuickBasisRef(mol.symbols, mol.xyz, "6-31G*")
    E_ref, _, _ = O.rhf(O.Oracle(qb), mol.nelec(), cutoff=1e-10)
    print(f"  served {res[0].energy:.10f} oracle {E_ref:.10f}", flush=True)
    assert abs(res[0].energy - E_ref) <= 1e-6
