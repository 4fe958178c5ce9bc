"""Serving loop: admission by HBM footprint, and every request renders the
same bytes alone and inside the continuous-batching mix."""
import time

import numpy as np
import pytest
import torch

from imagebasedspatiotemporalinterpolation_amd import serving as S

pytestmark = pytest.mark.gpu

SMALL = [S.SessionClass("320x180", 1, 320, 180), S.SessionClass("480x272", 1, 480, 272),
         S.SessionClass("stereo 258x276", 3, 258, 276, eyes=2)]


@pytest.fixture(scope="module")
def cache():
    return S.KeyframeCache(seed=3)


def test_admission_by_hbm_budget(cache):
    fp = S.session_footprint(320, 180)
    srv = S.Server(hbm_budget=int(fp * 1.5), cache=cache)
    a = srv.open_session(SMALL[0])
    b = srv.open_session(SMALL[0])
    assert srv.active(a) and not srv.active(b)  # the second does not fit yet
    srv.submit(S.Request(a, 1, time.perf_counter()))
    srv.drain()
    srv.close_session(a)
    assert srv.active(b)  # admitted once memory freed up
    with pytest.raises(ValueError):
        srv.open_session(S.SessionClass("huge", 1, 4096, 4096))
    srv.close()


@pytest.mark.parametrize("graph", [True, False])
def test_results_same_alone_and_in_mix(cache, graph):
    srv = S.Server(hbm_budget=1 << 34, max_batch=4, n_streams=2, graph=graph, cache=cache)
    sids = [srv.open_session(SMALL[i % 3]) for i in range(7)]
    rng = np.random.default_rng(0)
    reqs = []
    t0 = time.perf_counter()
    for k in range(40):  # interleaved arrivals; several per session, some back to back
        sid = sids[int(rng.integers(0, len(sids)))]
        s = srv.sessions[sid]
        outs = [torch.from_numpy(rng.integers(0, 256, (s.cls.height, s.cls.width, 4), dtype=np.uint8)).cuda()
                for _ in range(s.cls.eyes)]
        prev = [o.cpu().numpy() for o in outs]
        r = S.Request(sid, 1 + k % (s.fpa - 1), t0, outs=outs)
        reqs.append((r, prev))
        srv.submit(r)
        if k % 3 == 0:
            srv.step()
    srv.drain()
    assert srv.stats.batches >= 2 and max(srv.stats.batch_sizes) > 1
    for r, prev in reqs:
        s = srv.sessions[r.sid]
        for e, rend in enumerate(s.renderers):
            alone = torch.from_numpy(prev[e].copy()).cuda()
            rend.render(alone, s.info(r.b, e, srv.max_iter))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(r.outs[e].cpu().numpy(), alone.cpu().numpy(),
                                          err_msg=f"session {r.sid} frame {r.b} eye {e}")
    srv.close()
