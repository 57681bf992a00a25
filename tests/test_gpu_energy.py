"""Energy components of the e_kin.TAG / e_mag_oc.TAG lines (reference
``output.py:111-193``) from the device reduction (``mh_energy`` with per-part
bins) against the oracle on the same random fields."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(l_max=16, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35)


def test_energy_components_match_oracle():
    import torch
    from magic_amd.dynamo import DynamoRun
    from oracle.dynamo import OracleRun
    run, orc = DynamoRun(dict(CFG)), OracleRun(dict(CFG))
    rng = np.random.default_rng(5)
    L, N = run.f["w"].shape
    for k in ("w", "dw", "z", "b", "db", "aj"):
        a = rng.standard_normal((L, N)) + 1j * rng.standard_normal((L, N))
        getattr(orc, k)[:] = a
        run.f[k].copy_(torch.from_numpy(a))
    ek, em = run.energy_components()
    oek, oem = orc.energy_components()
    np.testing.assert_allclose(ek, oek, rtol=1e-12)
    np.testing.assert_allclose(em, oem, rtol=1e-12)
    e = run.energies()
    assert ek[0] == pytest.approx(e[0], rel=1e-13) and ek[1] == pytest.approx(e[1], rel=1e-13)
    assert em[0] == pytest.approx(e[2], rel=1e-13) and em[1] == pytest.approx(e[3], rel=1e-13)
    run.close()
