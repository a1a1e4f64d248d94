"""Host-side logic of the batched scan: the 64-pulse workload and the
validation of scan inputs (no GPU needed: errors are raised before the
library is asked to allocate anything)."""
import dataclasses

import numpy as np
import pytest

from sbewh_amd.scan import Scan, ellipticity_scan
from tests import cases


def test_ellipticity_scan_workload():
    cfgs = ellipticity_scan()
    assert len(cfgs) == 64
    assert sorted({c.num_k for c in cfgs}) == [3600, 8100, 14400, 22500]
    assert sum(c.num_k for c in cfgs) == 777600
    assert all(c.model is cfgs[0].model for c in cfgs)
    # parameters in the stated ranges, reproducible from the seed
    e1 = np.array([c.pulses[0].Emax_V_nm for c in cfgs])
    eps = np.array([c.pulses[1].Emax_V_nm / c.pulses[0].Emax_V_nm if len(c.pulses) > 1 else 0.0 for c in cfgs])
    E0 = e1 * np.sqrt(1 + eps ** 2)
    assert (E0 >= 0.3).all() and (E0 <= 1.5).all() and (eps <= 1).all()
    again = ellipticity_scan()
    assert [c.pulses[0].cep for c in again] == [c.pulses[0].cep for c in cfgs]


def test_scan_rejects_mixed_settings():
    a = cases.small_square()
    b = dataclasses.replace(a, eps_rel=1e-6)
    with pytest.raises(ValueError):
        Scan([a, b])
    c = dataclasses.replace(a, model=dataclasses.replace(a.model))
    with pytest.raises(ValueError):
        Scan([a, c])
    with pytest.raises(ValueError):
        Scan([])
