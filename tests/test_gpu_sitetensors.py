"""The site tensors fillsitetensors! builds (T_b = Pi1·P⁻¹, tensorci2.jl:367-394)
read back from the device before the final sweep1site! overwrites them
(TCI_FLAG_NO_FINAL_SWEEP), against the oracle's first optimize! iteration
(sweep2site! with fillsitetensors, tensorci2.jl:777-787), bit for bit."""
import numpy as np
import pytest

import oracle as O
from problems import c4_spec, specs
from test_gpu_solves import _tci_problem

pytestmark = pytest.mark.gpu

_NAMES = {"lorentz8", "lorentz5_tight", "gauss_sum", "weighted_lorentz", "random_tt", "quantics_gc",
          "gauss_sum_d2"}
_SPECS = [s for s in specs() if s[0] in _NAMES] + [c4_spec(0)]


@pytest.mark.parametrize("spec", _SPECS, ids=lambda s: s[0])
def test_fillsitetensors_on_device_equals_oracle(tci, ctx, spec):
    name, kind, params, ip, ld, piv, opts = spec
    print("sitetensors", name, flush=True)
    f, ld_, piv_, kw = _tci_problem(tci, spec)
    kw.update(maxiter=1, flags=1)
    (t,) = tci.crossinterpolate2_batch([(f, ld_, piv_, kw)])
    assert t.status == 0
    o = O.TCI(ld)
    of = O.Functor(kind, params, ip)
    o.init(of, piv if piv is not None else [[1] * len(ld)])
    ss = {"backandforth": 0, "forward": 1, "backward": 2}[opts.get("sweepstrategy", "backandforth")]
    o.sweep2site(of, niter=2, abstol=opts["tolerance"] * o.maxsamplevalue,
                 maxbonddim=opts.get("maxbonddim", O.INT64_MAX), sweepstrategy=ss,
                 strictlynested=opts.get("strictlynested", False), fillsitetensors=True)
    for b in range(len(ld)):
        assert np.array_equal(t.sitetensors[b], o.sitetensor(b)), (name, b)
