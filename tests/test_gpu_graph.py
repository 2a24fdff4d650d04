"""Stage graphs (vnlbhip_set_graph): a stage's launch sequence on both
streams is captured once after an eager run and replayed.  Every replay,
with eager work of another context between replays, must reproduce the
eager stage on the same inputs: the plain form up to the order of its fp64
accumulator adds, the fast form up to the order of its fp32 adds; the
group statistics come from the device summary and must match too.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip.params import BASIC, FINAL

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    return torch.from_numpy(a[:, None] if a.ndim == 3 else a).to("cuda:0")


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_graph_replays_match_eager(form):
    noisy, _, bp, fp = workloads.make("yuv_pt2_tau200")
    # Different content at the same address on every call: a replay that
    # missed part of the launch sequence (for instance the paste scans of the
    # side stream) would reuse stale results.
    clips = [_dev(noisy), _dev(np.ascontiguousarray(noisy[:, :, ::-1, ::-1]))]
    x = torch.empty_like(clips[0])
    eager = vnlbhip.Context(0)
    graph = vnlbhip.Context(0)
    tol = 1e-6 if form == "plain" else 1e-4
    try:
        eager.set_form(form)
        graph.set_form(form)
        graph.set_graph(True)
        out_b = torch.empty_like(x)
        out_f = torch.empty_like(x)
        for i in range(4):
            print(f"[graph] {form} call {i}", flush=True)
            x.copy_(clips[i % 2])
            graph.stage(BASIC, bp, x, out=out_b)
            stats_b = graph.stats()
            graph.stage(FINAL, fp, x, out_b, out=out_f)
            stats_f = graph.stats()
            want_b = eager.stage(BASIC, bp, x)
            eager_b = eager.stats()
            want_f = eager.stage(FINAL, fp, x, out_b.clone())
            eager_f = eager.stats()
            torch.cuda.synchronize()
            np.testing.assert_allclose(out_b.cpu().numpy(), want_b.cpu().numpy(), rtol=0, atol=tol)
            np.testing.assert_allclose(out_f.cpu().numpy(), want_f.cpu().numpy(), rtol=0, atol=tol)
            for key in ("anchors", "groups", "retained", "batches", "parts"):
                assert stats_b[key] == eager_b[key], (i, key)
                assert stats_f[key] == eager_f[key], (i, key)
        # The first call of each stage runs eagerly and captures; the later
        # ones replay.
        assert graph.stats()["graph_replays"] >= 4
    finally:
        eager.close()
        graph.close()
