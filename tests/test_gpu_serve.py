"""Continuous-batching SCF server: every request's energy is the same when
served alone and in a mix of concurrently running jobs (mixed precision,
incremental builds, QUICK's cutoff schedule; tolerance 1e-8 Eh), and within
1e-6 Eh of the FP64 SCF of the oracle."""
import numpy as np
import pytest

from tests.helpers import announce

pytestmark = pytest.mark.gpu


def test_served_alone_equals_served_in_the_mix():
    """24 Poisson arrivals (40/s) of molecules with 3-20 atoms in three bases,
    a budget that holds a few jobs at a time: prepared on the worker thread,
    admitted at regroups, retired into the running batch."""
    announce("serving: 24 Poisson requests, alone vs mix")
    import torch
    from quick_amd.serve import FockServer, Request, request_stream
    dev = torch.device("cuda:0")
    reqs = request_stream(24, 40.0, seed=11, max_atoms=20)
    budget = 3 * 32 * 2 ** 20
    srv = FockServer(dev, hbm_budget=budget, precision="mixed")
    mix, stats = srv.run(reqs)
    print(f"  {stats}", flush=True)
    assert len(mix) == len(reqs) and all(r.converged for r in mix)
    assert stats["rebuilds"] > 1
    mix = {r.rid: r for r in mix}
    worst = 0.0
    for r in reqs:
        alone, _ = FockServer(dev, precision="mixed").run([Request(r.rid, r.mol, r.basis, 0.0)])
        worst = max(worst, abs(alone[0].energy - mix[r.rid].energy))
    print(f"  max |alone - mix| = {worst:.1e} Eh", flush=True)
    assert worst <= 1e-8


@pytest.mark.parametrize("name,basis", [("water", "6-31G*"), ("alkane2", "def2-SVP")])
def test_served_energy_matches_oracle_scf(name, basis):
    announce(f"serving: {name} {basis} vs oracle SCF")
    import torch
    from oracle import oracle as O
    from quick_amd.serve import FockServer, Request
    from tests.helpers import molecule
    mol = molecule(name)
    res, _ = FockServer(torch.device("cuda:0"), precision="mixed").run([Request(0, mol, basis, 0.0)])
    qb = O.QuickBasisRef(mol.symbols, mol.xyz, basis)
    E_ref, _, _ = O.rhf(O.Oracle(qb), mol.nelec(), cutoff=1e-10)
    print(f"  served {res[0].energy:.10f} oracle {E_ref:.10f} (final cutoff {res[0].cutoff:.0e}, "
          f"{res[0].full_builds} full builds in {res[0].iterations})", flush=True)
    assert res[0].converged
    assert abs(res[0].energy - E_ref) <= 1e-6
    assert res[0].full_builds < res[0].iterations          # deltaO builds ran
