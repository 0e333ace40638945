"""Batched updates of many volumes (ddgi_batch_*) against per-volume ddgi_update calls.

Both paths run the same per-ray and per-probe device code, so the bar is bit-identical atlases,
states and ray records, for every blend mode, every work-item / tile configuration, and through
the autotuner (which must leave the volumes as it found them).  Five volumes over three small
scenes with uneven grids (4x4x4 .. 8x4x8) and rays per probe (64, 100, 128, 256, 512)."""
import numpy as np
import pytest

from vulkanexp_amd import IrradianceProbes, ProbeBatch, scenes
from vulkanexp_amd import _lib
from vulkanexp_amd._lib import DdgiError

pytestmark = pytest.mark.gpu

SPECS = [(0, (8, 4, 8), 128), (1, (4, 4, 4), 512), (2, (6, 4, 5), 100), (0, (5, 5, 5), 64), (1, (7, 3, 6), 256)]


@pytest.fixture(scope="module")
def scene_list():
    return [scenes.room_boxes_scene(1), scenes.room_boxes_scene(2, open_top=True),
            scenes.clustered_scene(3, n_instances=4)]


def make_volumes(scene_list, specs=SPECS):
    light = scenes.sun_light()
    vols = []
    for si, res, n in specs:
        lo, hi = scenes.inset_extent(scene_list[si])
        p = IrradianceProbes(scenes.grid_params(lo, hi, res, n), max_rays_per_probe=n)
        p.set_scene(scene_list[si])
        p.set_light(light)
        p.set_sky(scenes.SKY_RADIANCE)
        vols.append(p)
    for k, p in enumerate(vols):
        p.init_probes(scenes.orientations(100 + k, 1)[0])
    return vols


def frame_orientations(frame, count):
    return np.stack([scenes.orientations(1000 * frame + v, 1)[0] for v in range(count)])


def state_of(vols):
    for p in vols:
        p.synchronize()
    return [(p.irradiance(_lib.LAYOUT_TILED), p.depth(_lib.LAYOUT_TILED), p.states()) for p in vols]


def assert_same(a, b, what):
    for v, (x, y) in enumerate(zip(a, b)):
        for name, u, w in zip(("irradiance", "depth", "states"), x, y):
            assert np.array_equal(u, w), f"{what}: volume {v} {name} differs (max |d| {np.abs(u.astype(np.float64) - w).max():.3e})"


def destroy(*groups):
    for g in groups:
        for p in g:
            p.destroy()


@pytest.mark.parametrize("flags", [0, _lib.UPDATE_BLEND_BF16, _lib.UPDATE_BLEND_FP16, _lib.UPDATE_HYSTERESIS_RAMP],
                         ids=["fp32", "bf16", "fp16", "ramp"])
def test_batch_matches_per_volume_updates(scene_list, flags):
    seq, bat = make_volumes(scene_list), make_volumes(scene_list)
    batch = ProbeBatch(bat)
    try:
        for f in range(3):
            Rs = frame_orientations(f, len(SPECS))
            for p, R in zip(seq, Rs):
                p.update(R, _lib.UPDATE_ALL_PROBES | flags)
            batch.update(Rs, flags)
            assert_same(state_of(seq), state_of(bat), f"frame {f}")
        # ray and hit records of the last update too
        for v, (p, q) in enumerate(zip(seq, bat)):
            rp, ip = p.last_update_rays()
            rq, iq = q.last_update_rays()
            assert rp.shape[0] == p.get_probe_count() and np.array_equal(ip, iq), v
            assert np.array_equal(rp, rq), v
            assert np.array_equal(p.last_update_hits(), q.last_update_hits()), v
    finally:
        batch.destroy()
        destroy(seq, bat)


@pytest.mark.parametrize("cfg", [dict(trace_block=64, trace_item_blocks=1, blend_probes=4),
                                 dict(trace_block=256, trace_item_blocks=16, blend_probes=32),
                                 dict(trace_block=128, trace_item_blocks=64, blend_probes=8),
                                 dict(trace_block=256, trace_item_blocks=8, blend_probes=64)],
                         ids=["b64-i1-p4", "b256-i16-p32", "b128-i64-p8", "b256-i8-p64"])
def test_configurations_give_identical_results(scene_list, cfg):
    a, b = make_volumes(scene_list), make_volumes(scene_list)
    ba, bb = ProbeBatch(a), ProbeBatch(b, config=cfg)
    try:
        assert bb.config == cfg
        for f in range(2):
            Rs = frame_orientations(f, len(SPECS))
            ba.update(Rs)
            bb.update(Rs)
        assert_same(state_of(a), state_of(b), str(cfg))
    finally:
        ba.destroy()
        bb.destroy()
        destroy(a, b)


def test_autotune_restores_state_and_picks_a_measured_config(scene_list):
    vols = make_volumes(scene_list)
    batch = ProbeBatch(vols)
    try:
        Rs = frame_orientations(0, len(SPECS))
        batch.update(Rs, _lib.UPDATE_HYSTERESIS_RAMP)
        before = state_of(vols)
        hyst = [float(p.params()["hysteresis"]) for p in vols]
        best, records = batch.autotune(frame_orientations(1, len(SPECS)), _lib.UPDATE_HYSTERESIS_RAMP, iterations=2)
        assert len(records) == 9 + 5
        assert all(r["ms"] > 0 for r in records)
        assert best in [{k: r[k] for k in ("trace_block", "trace_item_blocks", "blend_probes")} for r in records]
        assert min(r["ms"] for r in records) == pytest.approx(
            min(r["ms"] for r in records if {k: r[k] for k in best} == best))
        assert batch.config == best
        assert_same(before, state_of(vols), "after autotune")
        assert hyst == [float(p.params()["hysteresis"]) for p in vols]
        print("autotune:", best, [(r["trace_block"], r["trace_item_blocks"], r["blend_probes"], round(r["ms"], 4)) for r in records])
    finally:
        batch.destroy()
        destroy(vols)


def test_timings_and_handle_lifetime(scene_list):
    vols = make_volumes(scene_list)
    batch = ProbeBatch(vols)
    try:
        batch.update(frame_orientations(0, len(SPECS)))
        total, per = batch.timings()
        print(f"batched update {total:.4f} ms, per-volume completion {np.round(per, 4)}")
        assert total > 0.0
        assert np.all(per > 0.0) and per.max() <= total + 0.5
        # a handle held by a batch cannot be destroyed
        with pytest.raises(DdgiError) as e:
            vols[0].destroy()
        assert e.value.status == 6
        # partial schedules and visit counting are not batched; bad arguments are refused
        import ctypes as C
        R = np.ascontiguousarray(frame_orientations(1, len(SPECS)), np.float32)
        Rp = R.ctypes.data_as(C.c_void_p)
        lib = batch._lib
        assert lib.ddgi_batch_update(batch.handle, Rp, 0, None) == 4
        assert lib.ddgi_batch_update(batch.handle, Rp, _lib.UPDATE_ALL_PROBES | _lib.UPDATE_COUNT_VISITS, None) == 4
        assert lib.ddgi_batch_update(batch.handle, None, _lib.UPDATE_ALL_PROBES, None) == 1
        assert lib.ddgi_batch_set_config(batch.handle, C.byref(_lib.BatchConfig(96, 1, 16))) == 1
        assert lib.ddgi_batch_set_config(batch.handle, C.byref(_lib.BatchConfig(128, 1, 12))) == 1
    finally:
        batch.destroy()
        destroy(vols)


def test_duplicate_handle_is_rejected(scene_list):
    vols = make_volumes(scene_list, SPECS[:2])
    try:
        with pytest.raises(DdgiError):
            ProbeBatch([vols[0], vols[1], vols[0]])
    finally:
        destroy(vols)
