"""GPU: the GI request server.

* Each result is the same alone and in the mix: a burst of requests of
  mixed scenes, resolutions, VPL counts and light poses (admitted together,
  batched) gives the same irradiance, bit for bit, as the same requests
  spaced out so that each runs alone.
* HBM-budget admission: with a budget that holds one large request at a
  time, every request still completes and at most one large one is in
  flight; a request larger than the whole budget is rejected.
* One served result against the oracle (mixed-form tolerance), with the
  request's inputs rebuilt on the host (the device producer's RSM equals
  the host renderer's).
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

RES = [(64, 64), (128, 96)]
VPLS = [256, 1024]


@pytest.fixture(scope="module")
def server(ctx):
    scenes = [(W._scene(1, 500), 20_000, 5), (W._scene(2, 2_000), 40_000, 6)]
    srv = mlgi.Server(ctx, scenes, VPLS, RES, hbm_budget=256 << 20, max_inflight=16, max_batch=8)
    yield srv
    srv.close()


def trace(spacing):
    reqs = []
    k = 0
    for sc in (0, 1):
        for (w, h) in RES:
            for n in VPLS:
                reqs.append((sc, w, h, n, 100 + k, k * spacing))
                k += 1
    return reqs


def test_same_alone_and_in_the_mix(server):
    burst = trace(0.0)
    outs_mix = [np.zeros((h, w, 3), np.float32) for _, w, h, _, _, _ in burst]
    st, lat = server.run(burst, outs_mix)
    print(f"[serve] burst: {st}")
    assert st["completed"] == len(burst) and st["rejected"] == 0
    assert st["mean_batch"] > 1.0  # the burst was batched
    alone = trace(0.05)
    outs_alone = [np.zeros((h, w, 3), np.float32) for _, w, h, _, _, _ in alone]
    st2, _ = server.run(alone, outs_alone)
    print(f"[serve] spaced: {st2}")
    assert st2["completed"] == len(alone)
    for i, (a, b) in enumerate(zip(outs_mix, outs_alone)):
        assert np.array_equal(a, b), i
        assert np.abs(a).sum() > 0, i


def test_served_result_against_oracle(server):
    sc, w, h, n, seed = 1, 128, 96, 1024, 4242
    out = np.zeros((h, w, 3), np.float32)
    st, _ = server.run([(sc, w, h, n, seed, 0.0)], [out])
    assert st["completed"] == 1
    scene = W._scene(2, 2_000)
    light = mlgi.Camera.light(seed)
    dif, nrm, dep, inv = scene.render_rsm(light, 128, 128)
    cam = mlgi.Camera.default(w / h)
    gd, gn = scene.render_gbuffer(cam, w, h)
    vp_inv = cam.matrices()[2]
    pts, off = scene.sample_points(40_000, n, n, 6)
    sh = mlgi.shuffle_indices(n)
    pos, nrmv, pk, col = oracle.vpl_extract(dif, nrm, dep, inv, 5.0, sh, n, np.float32)
    atlas_side = mlgi.ism_indices_1d(n) * 32
    atlas = oracle.ism_splat(pts, off, pos, nrmv, atlas_side, 50.0)
    r16 = oracle.pullpush(atlas, 32, 3, 50.0)
    ref, _ = oracle.final_gather(gd, gn, vp_inv, pos, nrmv, col, r16, atlas_side, 50.0, 3000.0, 1e-3, 4)
    err = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    print(f"[serve] rel-L2 vs oracle {err:.3e}")
    assert err <= 2e-3


def test_budget_admission_and_rejection(ctx):
    scenes = [(W._scene(1, 500), 20_000, 5)]
    # 1024 VPLs -> 1024^2 atlas (6 MiB of atlas + ISM); 2048 VPLs -> 4096^2 (96 MiB)
    srv = mlgi.Server(ctx, scenes, [1024, 2048], [(64, 64)], hbm_budget=40 << 20, max_inflight=8, max_batch=8)
    try:
        reqs = [(0, 64, 64, 1024, 7 + i, 0.0) for i in range(12)] + [(0, 64, 64, 2048, 99, 0.0)]
        st, lat = srv.run(reqs)
        print(f"[serve] budget: {st}")
        assert st["completed"] == 12 and st["rejected"] == 1
        assert lat[-1] < 0 and (lat[:12] > 0).all()
        assert st["peak_bytes"] <= 40 << 20
        assert st["batches"] >= 2  # 40 MiB holds fewer than 12 requests of ~6.3 MiB
    finally:
        srv.close()
