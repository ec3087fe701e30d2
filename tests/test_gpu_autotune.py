"""The kernel tuning (ekm_tuning) changes launch geometry and memory
placement only: every allowed value of every parameter gives the same state
bit for bit, on a stratified case with both subgrid models and the headline
projection form, and on the 64x65x65 case with the MUDPACK replica.  The
autotuner's search returns one of those tunings."""
import numpy as np
import pytest

from ekman_bl_amd import autotune as AT
from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import config, mg_factors

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "les_fp32":
        cfg = config(2)
        p = cfg.params
        p.nx, p.ny, p.nz = 32, 31, 31
        p.mg_ixp, p.mg_iex = mg_factors(p.nz + 2)
        p.mg_jyq, p.mg_jey = mg_factors(p.ny + 2)
        p.les_model, p.proj_mode = 2, 2
    else:
        cfg = config(1)
        p = cfg.params
        p.proj_mode = 0
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    return p, xp, yp, ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)


def _run(p, xp, yp, ic, tuning):
    s = S.DuctSolver(p, xp, yp, tuning=tuning)
    s.set_initial_physical(*ic)
    s.time_step(2)
    out = s.get_state()
    got = s.tuning().as_dict()
    s.close()
    return out, got


@pytest.mark.parametrize("name", ["les_fp32", "replica"])
def test_every_tuning_gives_the_same_state(name):
    p, xp, yp, ic = _case(name)
    ref, t0 = _run(p, xp, yp, ic, None)
    assert t0 == S.default_tuning().as_dict()
    for key in AT.relevant(p) + (["mg_block"] if p.proj_mode else ["tight_lds_kib"]):
        for val in AT.SPACE[key]:
            t = dict(t0, **{key: val})
            st, got = _run(p, xp, yp, ic, t)
            assert got == t
            for f, (a, b) in enumerate(zip(st, ref)):
                assert np.array_equal(a, b), (name, key, val, f)
            print(f"[tuning {name}] {key}={val}: identical", flush=True)


def test_autotuner_returns_an_allowed_tuning(tmp_path):
    p, xp, yp, ic = _case("les_fp32")
    res = AT.tune(p, xp, yp, ic, steps=1, log=lambda m: print(m, flush=True))
    for k, v in res["tuning"].items():
        assert v in AT.SPACE[k]
    assert res["ms_per_step"] <= res["trials"][0]["ms_per_step"]
    path = str(tmp_path / "tuning.json")
    AT.save("small", 2, res, path, device="dev")
    assert AT.load("small", 2, path, device="dev") == res["tuning"]
