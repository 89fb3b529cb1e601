"""Cameras and depth/normal prepasses for the view tier's tests.

A camera is a right-handed look-at view with a reverse-Z infinite
perspective (clip = (f/a x, f y, near, -z), depth = near / -z: the
convention of the prepass the view tier decodes, depth 0 = nothing). The
prepass is ray cast through the oracle's traversal: each pixel's ray from the
camera through its centre, the hit point's depth and the hit face's normal.
"""
import math

import numpy as np


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = r, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    return view


def perspective(fov_y, aspect, near=0.1):
    f = 1.0 / math.tan(fov_y / 2.0)
    p = np.zeros((4, 4))
    p[0, 0], p[1, 1], p[2, 3], p[3, 2] = f / aspect, f, near, -1.0
    return p


def camera(eye, target, width, height, fov_y=math.radians(60.0)):
    """(clip_from_world, world_from_clip) as float64 4x4 matrices."""
    clip_from_world = perspective(fov_y, width / height) @ look_at(eye, target)
    return clip_from_world, np.linalg.inv(clip_from_world)


def prepass(tracer, eye, clip_from_world, world_from_clip, width, height):
    """Depth [h, w] and normals [h, w, 3] (float32) by ray casting the oracle's scene."""
    depth = np.zeros((height, width), np.float32)
    normals = np.zeros((height, width, 3), np.float32)
    eye = np.asarray(eye, np.float64)
    for y in range(height):
        for x in range(width):
            ndc = np.array([(x + 0.5) / width * 2.0 - 1.0, 1.0 - (y + 0.5) / height * 2.0, 1.0, 1.0])
            p = world_from_clip @ ndc
            d = p[:3] / p[3] - eye
            d /= np.linalg.norm(d)
            oi, of = tracer.trace_ray(eye.astype(np.float32), d.astype(np.float32))
            if oi[4] != 1:
                continue
            hit = np.array([of[0], of[1], of[2], 1.0])
            c = clip_from_world @ hit
            depth[y, x] = np.float32(c[2] / c[3])
            normals[y, x] = of[3:6]
    return depth, normals
