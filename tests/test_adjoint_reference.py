"""The torch restatement used as the FP64 reference gradient (tests/
adjoint_reference.py) against the oracle's forward run: the same states, so
its autograd gradient is the derivative of the step the oracle pins."""
import numpy as np
import torch

import adjoint_reference as AR
import goldens
from gfsph import cases as C
from gfsph.dem import fractal_dem
from oracle import oracle


def small_case(integrator=2, n=250, mu=0.2, xi=500.0):
    dem = fractal_dem(65, 2.0, 17, relief=30.0)
    parts = C.elliptical_scar(40.0, 64.0, n, 1.0, 2.0, 5, h_range=(1.0, 3.0))
    prm = C.Params(integrator=integrator, dt=0.02, h_inf=1e-3, const=C.voellmy(mu, xi))
    return C.Case("adjref", prm, dem, parts, 0)


def test_torch_reference_matches_oracle():
    for integrator in (2, 4):
        case = small_case(integrator)
        nsteps = 30
        st = oracle.init(case)
        oracle.run(case, st, nsteps)
        c = case.params.cst()
        mass = torch.as_tensor(case.particles.mass, dtype=torch.float64)
        with torch.no_grad():
            x, v, h = AR.run(case, nsteps, torch.tensor(c[9], dtype=torch.float64),
                             torch.tensor(c[3], dtype=torch.float64), mass)
        assert np.all(st.flag == 0)
        assert goldens.rel_err(x.numpy(), st.x) <= 1e-12
        assert goldens.rel_err(v.numpy(), st.v) <= 1e-10
        assert goldens.rel_err(h.numpy(), st.h) <= 1e-12
