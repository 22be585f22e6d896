"""The no-allocation batch entry point (GPU), called through the C ABI: two calls
in a row -- four scenes of uneven sizes, then one smaller scene -- each give
exactly the fields and monitor sums of the allocating batch handle
(mw_batch_create / mw_batch_run) on the same scenes, and tables beyond the
static scratch are refused."""

import numpy as np
import pytest
import torch

from witwin_maxwell_amd import configs
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan

pytestmark = pytest.mark.gpu


def _runs(plans):
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    return [DeviceRun(p, "cuda") for p in plans]


def _same(a_runs, b_runs):
    for ra, rb in zip(a_runs, b_runs):
        for c in COMPONENTS:
            assert torch.equal(ra.fields[c], rb.fields[c]), c
        for x, y in zip(ra.plane_re + ra.plane_im, rb.plane_re + rb.plane_im):
            assert torch.equal(x, y)
        if ra.point_re is not None:
            assert torch.equal(ra.point_re, rb.point_re) and torch.equal(ra.point_im, rb.point_im)
        assert float(ra.fields["Ez"].abs().max()) > 0


def test_noalloc_batch_matches_batch_handle():
    import test_gpu_batch as tb
    from witwin_maxwell_amd.fdtd.batch import BatchRun, run_batch_noalloc
    print("case: no-allocation batch, 4 scenes then 1 smaller scene, 80 steps each", flush=True)
    scenes = tb._scenes()
    plans = [compile_plan(s, **dict(kw, time_steps=80)) for s, kw in scenes]
    small, kw = configs.vacuum_dipole(cells=10, pml=4)
    small_plan = compile_plan(small, **dict(kw, time_steps=80, full_field_dft=False))
    for group in (plans, [small_plan]):
        ref = BatchRun(_runs(group))
        ref.run(80)
        got = _runs(group)
        run_batch_noalloc(got, 50)
        run_batch_noalloc(got, 30)
        torch.cuda.synchronize()
        _same(ref.runs, got)
        ref.close()


def test_noalloc_batch_refuses_tables_beyond_scratch():
    from witwin_maxwell_amd.fdtd.batch import run_batch_noalloc
    print("case: no-allocation batch, 200 scenes (tables beyond the scratch)", flush=True)
    scene, kw = configs.vacuum_dipole(cells=6, pml=3)
    plan = compile_plan(scene, **dict(kw, time_steps=4, full_field_dft=False))
    runs = _runs([plan] * 200)
    with pytest.raises(RuntimeError, match="scratch"):
        run_batch_noalloc(runs, 2)
