"""ddgi_reflection_noalloc (the reflection pass without device allocations, scratch in a static
device array) against ddgi_reflection on the same inputs, through the C-ABI: bit-identical outputs
on two calls in a row (the second with fewer texels, so stale scratch contents from the first call
must not leak into it), no change of the device's free memory across the calls, and the texel limit
refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
from vulkanexp_amd import _lib
from vulkanexp_amd._lib import DdgiError

pytestmark = pytest.mark.gpu

W, H = 192, 108
EYE = (2.0, 7.0, 2.0)


def _free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.fixture(scope="module")
def setup():
    import torch
    scene, grid = scenes.config_small(seed=1, open_top=True)
    p = IrradianceProbes(grid, max_rays_per_probe=128)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    for R in scenes.orientations(61, 3):
        p.update(R, _lib.UPDATE_ALL_PROBES)
    rays = gbuffer.camera_rays(W, H, eye=EYE, look_at=(12.0, 1.0, 12.0))
    gb = p.gbuffer(torch.from_numpy(rays).cuda())
    gb[:, 11] = 0.25  # every hit texel glossy: each one traces and shades a reflected ray
    gen = torch.Generator(device="cuda").manual_seed(3)
    noise = torch.rand((W * H, 2), device="cuda", generator=gen)
    torch.cuda.synchronize()
    yield p, gb, noise
    p.destroy()


@pytest.mark.parametrize("jitter", [False, True])
def test_noalloc_matches_reflection_on_two_calls(setup, jitter):
    import torch
    p, gb, noise = setup
    nz = noise if jitter else None
    sizes = (W * H, W * H // 3 + 17)  # the second call is smaller
    inputs = [(gb[:n].contiguous(), None if nz is None else nz[:n].contiguous()) for n in sizes]
    wants = [p.reflection(g, EYE, noise=z) for g, z in inputs]
    outs = [torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda") for n in sizes]
    torch.cuda.synchronize()
    free0 = _free_device_bytes()
    for (g, z), got in zip(inputs, outs):
        p.reflection_noalloc(g, EYE, got, noise=z)
    torch.cuda.synchronize()
    assert _free_device_bytes() == free0, "device memory changed across the allocation-free calls"
    for n, got, want in zip(sizes, outs, wants):
        a, b = got.cpu().numpy(), want.cpu().numpy()
        traced = (a[:, :3] != 0).any(1).mean()
        print(f"{n} texels ({'jittered' if jitter else 'mirror'}): {traced:.3f} with reflected light")
        assert traced > 0.2
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{np.sum(a != b)} values differ at {n} texels"


def test_noalloc_refuses_more_texels_than_its_scratch(setup):
    import torch
    p, gb, _ = setup
    out = torch.empty((4, 4), dtype=torch.float32, device="cuda")
    lib = _lib.load()
    origin = np.asarray(EYE, np.float32)
    st = lib.ddgi_reflection_noalloc(p.handle, C.c_void_p(gb.data_ptr()), (1 << 21) + 1,
                                     origin.ctypes.data_as(C.c_void_p), None, C.c_void_p(out.data_ptr()), None)
    assert st == 1
    assert b"2,097,152" in lib.ddgi_last_error(p.handle)
    with pytest.raises(DdgiError):
        _lib.check(st, p.handle)
