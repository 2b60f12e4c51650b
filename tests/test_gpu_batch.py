"""GPU parity of the batched uneven workload (sp_plan_powder_batch): one
persistent launch over many spin systems of different reduced dimension."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, workloads

pytestmark = pytest.mark.gpu


def test_batch_matches_oracle_per_system():
    b = workloads.batch_uneven(seed=5, n_systems=10, nsteps=256, min_or=3, max_or=8, min_spins=2, max_spins=7,
                               workers=1)
    plan = _native.plan_powder_batch(b['systems'], b['dt'], b['nsteps'])
    st = plan.execute()
    fids = plan.result_batch()
    print('batch stats', st, flush=True)
    worst = 0.0
    for y, f in zip(b['systems'], fids):
        ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], b['dt'], b['nsteps'], variant=0)
        e = O.rel_err(f, ref)
        worst = max(worst, e)
    print(f'batch of 10 systems: worst rel err vs oracle {worst:.2e}', flush=True)
    assert worst < 1e-10


def test_batch_equals_single_system_calls():
    b = workloads.batch_uneven(seed=6, n_systems=4, nsteps=128, min_or=2, max_or=5, min_spins=2, max_spins=6,
                               workers=1)
    plan = _native.plan_powder_batch(b['systems'], b['dt'], b['nsteps'])
    plan.execute()
    fids = plan.result_batch()
    for y, f in zip(b['systems'], fids):
        g, _, _ = _native.powder_fid(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], b['dt'], b['nsteps'])
        assert O.rel_err(f, g) < 1e-12
