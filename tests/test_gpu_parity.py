"""GPU parity: every stage through the C ABI against the oracle.

Bars (stage-wise: each oracle stage consumes the GPU's previous stage):
  * VPL records: bit-exact vs the FP32 twin (positions, packed and unpacked
    normals, flux and its halves, lookAtRH basis); positions within 1e-5 of
    FP64.
  * ISM atlas: bit-exact vs the FP32 twin (integer atomicMin is order-free).
  * Pull-push R16: bit-exact vs the FP32 twin.
  * Irradiance: rel-L2 <= 1e-4 vs the FP64 transcription (plain FP32 form),
    <= 2e-3 (mixed form), with the mixed form's error printed per output
    channel and as per-pixel quantiles.
"""
import numpy as np
import pytest

import mlgi
import oracle
from mlgi import workloads as W

pytestmark = pytest.mark.gpu

REL_L2_FP32 = 1e-4
REL_L2_MIXED = 2e-3


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


class Frame:
    """Runs one frame stage by stage and keeps host copies of every output."""

    def __init__(self, ctx, wl, **pkw):
        self.wl = wl
        self.inp = inp = W.build_inputs(wl)
        self.params = p = wl.params(**pkw)
        self.stage = st = mlgi.GIStage(ctx, p, wl.n_vpl, wl.width, wl.height)
        self.rsm = mlgi.rsm_desc(ctx.upload(inp.rsm_diffuse), ctx.upload(inp.rsm_normal),
                                 ctx.upload(inp.rsm_depth), inp.inv_bvp)
        self.pts = mlgi.points_desc(ctx.upload(inp.points), ctx.upload(inp.bin_offsets))
        self.gb = mlgi.gbuffer_desc(ctx.upload(inp.gb_depth), ctx.upload(inp.gb_normal), inp.vp_inv)
        print(f"[frame] {wl.name} {wl.width}x{wl.height} N={wl.n_vpl} I={wl.interleave} "
              f"P={wl.n_points} params={pkw}", flush=True)
        st.vpl_processor.process(self.rsm, p.light_intensity, bool(p.shuffle_lights), p)
        self.rec = st.vpl_processor.host_records()
        st.ism.render(self.pts, st.vpl_processor, p)
        self.atlas = st.ism.softrender.numpy()
        st.ism.pullpush(p)
        self.r16 = st.ism.pushpull_result.numpy()
        st.compute_final_gathering(self.gb)
        self.img = st.gi_buffer.numpy()

    @property
    def vpl(self):
        r = self.rec
        return r[:, 0:3], r[:, 4:7], r[:, 16:19]

    def reference(self):
        """FP64 irradiance and lit-pair count of the GPU's VPLs and ISM (computed once)."""
        if getattr(self, "_ref", None) is None:
            pos, nrm, col = self.vpl
            p = self.params
            self._ref = oracle.final_gather(self.inp.gb_depth, self.inp.gb_normal, self.inp.vp_inv, pos, nrm, col,
                                            self.r16, self.wl.atlas_side, p.z_far, p.gi_intensity_factor,
                                            p.vpl_clamping_value, self.wl.interleave)
        return self._ref


@pytest.fixture(scope="module", params=["tiny", "small", "1080p"])
def frame(request, ctx):
    return Frame(ctx, W.WORKLOADS[request.param])


def test_vpl_extract_bit_exact(frame):
    inp, wl = frame.inp, frame.wl
    sh = mlgi.shuffle_indices(wl.n_vpl)
    pos, nrm, pk, col = oracle.vpl_extract(inp.rsm_diffuse, inp.rsm_normal, inp.rsm_depth, inp.inv_bvp,
                                           frame.params.light_intensity, sh, wl.n_vpl, np.float32)
    r = frame.rec
    assert np.array_equal(r[:, 0:3], pos)
    assert np.array_equal(r[:, 19].view(np.uint32), pk)
    assert np.array_equal(r[:, 4:7], nrm)
    assert np.array_equal(r[:, 16:19], col)
    # the flux as halves (slots 11 and 15), round to nearest even
    h = col.astype(np.float16).view(np.uint16).astype(np.uint32)
    assert np.array_equal(r[:, 11].view(np.uint32), h[:, 0] | (h[:, 1] << 16))
    assert np.array_equal(r[:, 15].view(np.uint32), h[:, 2])
    # zero-flux flags: bit 0 floats all zero, bit 1 halves all zero
    flags = (col == 0).all(axis=1).astype(np.uint32) | ((h == 0).all(axis=1).astype(np.uint32) << 1)
    assert np.array_equal(r[:, 7].view(np.uint32), flags)
    assert (flags != 0).any() and (flags == 0).any()
    f = mlgi.vpl_fields(r)
    n, x = f["normal"], f["position"]
    plane = n[:, 0] * x[:, 0] + n[:, 1] * x[:, 1] + n[:, 2] * x[:, 2]  # float32, unfused
    assert np.array_equal(f["plane"], plane)
    for i in range(0, wl.n_vpl, max(1, wl.n_vpl // 64)):
        s, u = oracle.look_at_rh(r[i, 4:7], np.float32)
        assert np.array_equal(r[i, 8:11], s) and np.array_equal(r[i, 12:15], u), i
    pos64, _, pk64, _ = oracle.vpl_extract(inp.rsm_diffuse, inp.rsm_normal, inp.rsm_depth, inp.inv_bvp,
                                           frame.params.light_intensity, sh, wl.n_vpl, np.float64)
    assert np.abs(r[:, 0:3] - pos64).max() <= 1e-5 * 10.0


def test_ism_splat_bit_exact(frame):
    pos, nrm, _ = frame.vpl
    ref = oracle.ism_splat(frame.inp.points, frame.inp.bin_offsets, pos, nrm, frame.wl.atlas_side,
                           frame.params.z_far)
    assert (ref != 0xFFFFFFFF).any()
    mism = int((frame.atlas != ref).sum())
    assert mism == 0, f"{mism} atlas texels differ"


def test_pullpush_bit_exact(frame):
    ref = oracle.pullpush(frame.atlas, frame.wl.ism_side, frame.wl.levels, frame.params.z_far)
    bad = np.argwhere(frame.r16 != ref)
    where = [(int(y), int(x), int(frame.r16[y, x]), int(ref[y, x])) for y, x in bad[:8]]
    assert len(bad) == 0, f"{len(bad)} R16 texels differ (row, col, gpu, oracle): {where}"


def test_final_gather_fp32_vs_fp64(frame):
    ref, lit = frame.reference()
    assert lit > 0
    err = rel_l2(frame.img, ref)
    print(f"[fg] {frame.wl.name} rel-L2 {err:.3e} lit pairs {lit}")
    assert err <= REL_L2_FP32


def test_final_gather_mixed_vs_fp64(frame, ctx):
    st = frame.stage
    st.params.precision = mlgi.PRECISION_MIXED
    try:
        st.compute_final_gathering(frame.gb)
        ctx.sync()
        img = st.gi_buffer.numpy()
    finally:
        st.params.precision = mlgi.PRECISION_FP32
    ref, _ = frame.reference()
    err = rel_l2(img, ref)
    per_channel = [rel_l2(img[..., c], ref[..., c]) for c in range(3)]
    pix = np.linalg.norm(img - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)
    q = np.quantile(pix[np.linalg.norm(ref, axis=-1) > 0], [0.5, 0.99, 0.999])
    print(f"[fg mixed] {frame.wl.name} rel-L2 {err:.3e} per channel {['%.3e' % e for e in per_channel]} "
          f"per-pixel rel err p50/p99/p99.9 {q[0]:.2e}/{q[1]:.2e}/{q[2]:.2e}")
    assert err <= REL_L2_MIXED
    assert max(per_channel) <= REL_L2_MIXED
    assert not np.array_equal(img, frame.img)  # the mixed kernel did run


@pytest.mark.parametrize("name,prec,bar", [("fp32", mlgi.PRECISION_FP32, REL_L2_FP32),
                                            ("mixed", mlgi.PRECISION_MIXED, REL_L2_MIXED)])
def test_batched_path_vs_fp64(frame, ctx, name, prec, bar):
    """The batched kernels (the benchmark's headline path) on the one frame:
    VPLs, atlas and ISM bit-identical to the stage-wise path, irradiance
    within the precision's bar of the FP64 oracle."""
    st = frame.stage
    st.params.precision = prec
    try:
        b = mlgi.Batch(ctx, [(st, frame.rsm, frame.pts, frame.gb)])
        st.gi_buffer.copy_from(np.full((frame.wl.height, frame.wl.width, 3), -1.0, np.float32))
        b.run()
        ctx.sync()
        b.close()
        rec, atlas, r16, img = (st.vpl_processor.host_records(), st.ism.softrender.numpy(),
                                st.ism.pushpull_result.numpy(), st.gi_buffer.numpy())
    finally:
        st.params.precision = mlgi.PRECISION_FP32
    assert np.array_equal(rec, frame.rec)
    assert np.array_equal(atlas, frame.atlas)
    assert np.array_equal(r16, frame.r16)
    ref, _ = frame.reference()
    err = rel_l2(img, ref)
    per_channel = [rel_l2(img[..., c], ref[..., c]) for c in range(3)]
    print(f"[fg batched {name}] {frame.wl.name} rel-L2 {err:.3e} per channel "
          f"{['%.3e' % e for e in per_channel]}")
    assert err <= bar and max(per_channel) <= bar


def test_frame_matches_stages(frame, ctx):
    st = frame.stage
    st.process(frame.rsm, frame.pts, frame.gb)
    ctx.sync()
    assert np.array_equal(st.ism.softrender.numpy(), frame.atlas)
    assert np.array_equal(st.ism.pushpull_result.numpy(), frame.r16)
    assert np.array_equal(st.gi_buffer.numpy(), frame.img)


def test_count_lit_pairs(frame):
    pos, nrm, col = frame.vpl
    p = frame.params
    _, lit = oracle.final_gather(frame.inp.gb_depth, frame.inp.gb_normal, frame.inp.vp_inv, pos, nrm, col,
                                 None, frame.wl.atlas_side, p.z_far, p.gi_intensity_factor,
                                 p.vpl_clamping_value, frame.wl.interleave, shadowing=False)
    got = frame.stage.count_lit_pairs(frame.gb)
    assert abs(got - lit) <= max(2, 1e-5 * lit), (got, lit)
