"""GPU: the batched workload at its specified size.

64 independent frames of the 40,012-triangle room in one batch: square
G-buffers of 128^2..1024^2 and 256..4,096 VPLs drawn at random (ISMs of
32^2 up to 1,024 VPLs, else 64^2; 3 or 4 pull-push levels), each with its
own light pose and device-rendered RSM, 1M points per VPL count, mixed
precision, interleave and gather tile picked per shape by the autotuner,
the whole batch captured in one HIP graph.

* Graph replays reproduce the eager batch bit for bit.
* Four frames of different shapes (both ISM sides, the smallest and largest
  resolution) are checked against the oracle on inputs rebuilt on the host:
  VPL records, atlas and unorm16 ISM bit-exact, irradiance within the mixed
  form's rel-L2 2e-3 of the FP64 transcription.
* Each of those frames, run alone as a batch of one, gives the same bits as
  in the batch of 64.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu


def outputs(st):
    return (st.vpl_processor.host_records(), st.ism.softrender.numpy(), st.ism.pushpull_result.numpy(),
            st.gi_buffer.numpy())


def clobber(st):
    st.gi_buffer.copy_from(np.full(st.gi_buffer.shape, -1.0, np.float32))
    st.ism.pushpull_result.copy_from(np.zeros(st.ism.pushpull_result.shape, np.uint16))


def pick_frames(specs):
    """Four frames of different shapes: the smallest and the largest
    resolution, and both ISM sides."""
    chosen = []
    for want in (lambda s: s.size == 128 and s.ism_side == 32, lambda s: s.size == 1024 and s.ism_side == 64,
                 lambda s: s.size in (256, 512) and s.ism_side == 64, lambda s: s.size in (256, 512) and s.ism_side == 32):
        for i, s in enumerate(specs):
            if want(s) and i not in chosen:
                chosen.append(i)
                break
    return chosen


@pytest.fixture(scope="module")
def b64(ctx):
    b = W.Batch64(ctx)
    shapes = sorted({(s.size, s.n_vpl) for s in b.specs})
    print(f"[batch64] {len(b.frames)} frames, {len(shapes)} shapes: {shapes}", flush=True)
    yield b
    b.close()


def test_batch64_graph_oracle_and_alone(ctx, b64):
    batch = mlgi.Batch(ctx, b64.frames)
    batch.run()
    ctx.sync()
    eager = [outputs(f[0]) for f in b64.frames]
    g = batch.capture()
    for rep in range(2):
        for f in b64.frames:
            clobber(f[0])
        g.launch()
        ctx.sync()
        for i, f in enumerate(b64.frames):
            for x, y in zip(outputs(f[0]), eager[i]):
                assert np.array_equal(x, y), (rep, i)
    g.close()
    batch.close()

    chosen = pick_frames(b64.specs)
    assert len(chosen) == 4, chosen
    for i in chosen:
        sp = b64.specs[i]
        st = b64.frames[i][0]
        p = st.params
        print(f"[batch64] frame {i}: {sp.size}^2, {sp.n_vpl} VPLs, ISM {sp.ism_side}, levels {sp.levels}, "
              f"tuned {b64.tune[i]}", flush=True)
        rec, atlas, r16, img = eager[i]
        h = b64.host_inputs(i)
        dif, nrm, dep, inv = h["rsm"]
        sh = mlgi.shuffle_indices(sp.n_vpl)
        pos, vn, pk, col = oracle.vpl_extract(dif, nrm, dep, inv, p.light_intensity, sh, sp.n_vpl, np.float32)
        assert np.array_equal(rec[:, 0:3], pos) and np.array_equal(rec[:, 16:19], col), i
        assert np.array_equal(rec[:, 19].view(np.uint32), pk), i
        assert np.array_equal(atlas, oracle.ism_splat(h["points"], h["bin_offsets"], pos, vn, p.atlas_side, p.z_far,
                                                      threads=16)), i
        assert np.array_equal(r16, oracle.pullpush(atlas, sp.ism_side, sp.levels, p.z_far, threads=16)), i
        il = p.interleave if p.use_interleaving else 1
        ref, lit = oracle.final_gather(h["gb_depth"], h["gb_normal"], h["vp_inv"], pos, vn, col, r16, p.atlas_side,
                                       p.z_far, p.gi_intensity_factor, p.vpl_clamping_value, il, threads=16)
        err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
        print(f"[batch64] frame {i}: rel-L2 {err:.3e} (lit pairs {lit})", flush=True)
        assert err <= 2e-3, (i, err)
        # alone, as a batch of one
        clobber(st)
        one = mlgi.Batch(ctx, [b64.frames[i]])
        one.run()
        ctx.sync()
        for x, y in zip(outputs(st), eager[i]):
            assert np.array_equal(x, y), i
        one.close()
