"""Interleaved scheduling of the one-wavefront batched kernel (qt_mcwf_params::sched_chunk): a launch
cut into chunks of stream intervals, every trajectory's chunk c before any chunk c + 1, gives the
same trajectories bit for bit as whole-trajectory launches, and the same as the oracle."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import JC60, PLM20, P1M1, coherent, ground, p1m1_initial

pytestmark = pytest.mark.gpu


def _run(spec, psi0, n, iv_list, chunk, prec=Q.QT_FP32_DERIV, controller="odeint", draws=None, graph=False,
         noalloc=False, seed=5):
    sq = Q.System(spec)
    ens = Q.Ensemble(sq, n, psi0, Q.mcwf_pars(seed=seed, precision=prec, controller=controller, sched_chunk=chunk,
                                              jump_log_cap=256))
    if draws is not None:
        ens.set_draws(draws)
    if graph:
        ens.set_graph(True)
    rows = []
    for iv in iv_list:
        rows.append(ens.advance_noalloc(iv) if noalloc else ens.advance(iv))
    per, _ = ens.counters()
    st, psi, jt, jc = ens.get_states()
    return np.concatenate(rows), per, psi, jt, jc, np.array([s.t for s in st])


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("chunk", [1, 3, 7])
@pytest.mark.parametrize("prec", [Q.QT_FP64, Q.QT_FP32_DERIV])
def test_forced_chunks_equal_whole_trajectories(chunk, prec):
    print(f"[gpu] sched: JC-60 96 traj, chunk {chunk}, precision {prec}", flush=True)
    iv = [np.full(12, 0.1), np.full(5, 0.2)]
    whole = _run(JC60, ground(60), 96, iv, -1, prec)
    chunked = _run(JC60, ground(60), 96, iv, chunk, prec)
    _same(whole, chunked)


def test_forced_chunks_match_oracle():
    print("[gpu] sched: JC-60 24 traj, chunk 2 vs oracle", flush=True)
    iv = np.full(15, 0.1)
    rows, per, psi, jt, jc, t = _run(JC60, ground(60), 24, [iv], 2, Q.QT_FP64, seed=9)
    so = O.System(JC60)
    for k in (0, 11, 23):
        r = O.mcwf_run(so, ground(60), O.params(rng="philox", seed=9, stream=k), O.run_spec("NDt", Dt=0.1, NDt=15))
        err = np.linalg.norm(psi[k] - r["psi"]) / np.linalg.norm(r["psi"])
        assert err <= 1e-9, (k, err)
        nj = int(r["counters"][7])
        np.testing.assert_array_equal(jc[k, :nj], r["jump_ch"])
        for c in (0, 1, 2, 4, 5, 7):
            assert per[k, c] == r["counters"][c], (k, c)


def test_automatic_chunks_beyond_resident_wavefronts():
    # more trajectories than the GPU holds wavefronts: the automatic schedule interleaves chunks
    print("[gpu] sched: JC-60 12000 traj T=2, automatic vs whole", flush=True)
    iv = [np.full(20, 0.1)]
    whole = _run(JC60, ground(60), 12000, iv, -1)
    auto = _run(JC60, ground(60), 12000, iv, 0)
    _same(whole, auto)


@pytest.mark.parametrize("mode", ["graph", "noalloc", "host_draws", "gsl", "interaction_picture"])
def test_chunk_paths(mode):
    print(f"[gpu] sched: chunked path {mode}", flush=True)
    if mode == "interaction_picture":
        spec, psi0, n = P1M1, p1m1_initial(), 64
    else:
        spec, psi0, n = PLM20, coherent(0.5, 20), 64
    iv = [np.full(9, 0.1), np.full(9, 0.1)]
    kw = {}
    if mode == "graph":
        kw["graph"] = True
    elif mode == "noalloc":
        kw["noalloc"] = True
    elif mode == "host_draws":
        kw["draws"] = np.random.default_rng(3).random((n, 4096))
    elif mode == "gsl":
        kw["controller"] = "gsl"
    whole = _run(spec, psi0, n, iv, -1, **kw)
    chunked = _run(spec, psi0, n, iv, 2, **kw)
    _same(whole, chunked)
