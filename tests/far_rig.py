"""The far tier's test rig (gpu/view/pixel_tests.rs:24-330), for the oracle
and for the GPU tracer alike.

A 32^3 page with a 32x32 floor (y 0..3, material 2), optionally an emissive
slab over its middle (x, z 8..23, y 28..29), seen straight down from 96
voxels above by a 16x16 prepass: pixel x -> world x = 2 px + 1, pixel y ->
world z = 2 py + 1 (two voxels per pixel), every depth 0.5 (the floor's top,
y = 4), every normal +y. At 4 prepass pixels per unit at view depth 1, a
voxel is 0.04 pixels across, so every floor face is a far face.
"""
import numpy as np

import oracle_lib as orc

AIR = 256
SIZE = 16
SKY = (0.16, 0.24, 0.38)  # VoxelGiEnvironment::default (sources.rs:19-23)
EYE = (16.0, 100.0, 16.0)


def view_at(shift=0.0):
    """world_from_clip (row-major) of the rig's camera, panned `shift` voxels along +x."""
    return np.array([[16.0, 0.0, 0.0, 16.0 + shift],
                     [0.0, 0.0, 8.0, 0.0],
                     [0.0, -16.0, 0.0, 16.0],
                     [0.0, 0.0, 0.0, 1.0]])


def grid(ceiling=None):
    g = np.full((32, 32, 32), AIR, np.uint16)
    g[:, 0:4, :] = 2
    if ceiling is not None:
        g[8:24, 28:30, 8:24] = ceiling
    return g


def optics(ceiling=None, custom_light=None, environment=SKY):
    """(diffuse, emission, lights (n, 12), environment)."""
    diffuse = np.full((256, 3), 0.45, np.float32)
    emission = np.zeros((256, 3), np.float32)
    if ceiling is not None:
        emission[ceiling] = 8.0
    lights = np.zeros((0, 12), np.float32)
    if custom_light is not None:
        # A point light Bevy does not draw (intensity 0, range 64) whose
        # direct light the presentation shows: custom visible rgb 100.
        light = np.zeros(12, np.float32)
        light[0:3] = custom_light
        light[3] = 64.0
        light[9:12] = 100.0
        lights = light.reshape(1, 12)
    return diffuse, emission, lights, environment


def prepass():
    depth = np.full((SIZE, SIZE), 0.5, np.float32)
    normals = np.zeros((SIZE, SIZE, 3), np.float32)
    normals[..., 1] = 1.0
    return depth, normals


def camera(vgi, frame, shift=0.0, samples=1):
    wfc = view_at(shift)
    return vgi.view_camera(wfc, EYE, SIZE, SIZE, samples=samples, clip_from_world=np.linalg.inv(wfc), forward=(0.0, -1.0, 0.0),
                           pixels_per_unit=4.0, frame=frame)


def settings(vgi, cell, rate, steps=128, voxel_pixels=4.0, history=32.0):
    return vgi.far_settings(voxel_pixels=voxel_pixels, measure=0, cell=cell, history=history, rate=rate, steps=steps)


def oracle(vgi, cell, rate, ceiling=None, custom_light=None, environment=SKY, reach=None, steps=128, far=True):
    """An oracle tracer with the rig's scene and far tier."""
    diffuse, emission, lights, env = optics(ceiling, custom_light, environment)
    scene = orc.Scene()
    scene.load_pages(*vgi.pages_of(grid(ceiling)))
    t = orc.Tracer(scene)
    t.set_materials(diffuse, emission)
    t.set_lights(lights)
    t.set_transport(bounces=3, samples=1, environment=env, limited_reach=reach is not None,
                    reach=768.0 if reach is None else reach)
    t.view_far(settings(vgi, cell, rate, steps) if far else None)
    return t


def gpu(vgi, cell, rate, ceiling=None, custom_light=None, environment=SKY, reach=None, steps=128, far=True):
    diffuse, emission, lights, env = optics(ceiling, custom_light, environment)
    g = vgi.GpuTracer(0)
    g.upload_pages(*vgi.pages_of(grid(ceiling)))
    g.set_materials(diffuse, emission)
    g.set_lights(lights)
    g.set_transport(bounces=3, samples=1, environment=env, limited_reach=reach is not None, reach=768.0 if reach is None else reach)
    g.view_far(settings(vgi, cell, rate, steps) if far else None)
    return g


def cells_of(size):
    return (SIZE + size - 1) // size


def f16(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def counts(output, cells, which=None):
    """Sample counts of the output's cells `which` (default all), checking face and stamp."""
    words = output[16:16 + 4 * cells].reshape(cells, 4)
    if which is not None:
        words = words[which]
    assert (words[:, 2] & 7 == 2).all(), "owned top faces"
    assert (words[:, 2] >> 3 == output[15]).all(), "stamped with this frame"
    return f16(words[:, 1] >> 16)


def interior(cells_x):
    """Cells wholly inside the 16-pixel view."""
    return [i for i in range(cells_x * cells_x) if i % cells_x < cells_x - 1 and i // cells_x < cells_x - 1]
