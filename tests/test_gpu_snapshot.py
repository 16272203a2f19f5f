"""big-wave's end-to-end golden vector (big-wave/test.py:853-960) through the
product path: sim dir written by ravesim.simdir.setup_simulation with its own
delta/beta (ravesim.xray, the default), run by rs_run_simdir on the GPU."""
import json

import numpy as np
import pytest

from oracle import bigwave as bw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,atol", [("c16", 1e-14), ("c8", None)])
def test_golden_vector_on_gpu(golden_dir, tmp_path, dtype, atol):
    import ravesim
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import setup_simulation
    d = json.loads((golden_dir / "bigwave_snapshot.json").read_text())
    cfg = dict(d["known_config"])
    cfg["dtype"] = dtype
    sd = setup_simulation(cfg, tmp_path / "sim")
    ctx = ravesim.Context(0)
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)[0]
    exp = np.array(d["known_data"])
    if atol is not None:
        # the reference's own bar (~1e-6 of the values) and the fixture's
        # relative limit (9 significant digits, ~1e-8); the tight pin is the
        # oracle on the same sim dir
        assert np.allclose(got, exp, rtol=0, atol=atol)
        assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < 5e-8
        ref = bw.run_simdir(sd, 0, dtype=np.complex128)[0]
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10
    else:
        assert np.linalg.norm(got - exp) / np.linalg.norm(exp) < 1e-4
    ctx.close()
