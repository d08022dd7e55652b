"""The light tree reference test's worlds, regenerated from their seeds (tests/light_tree_test.cpp:
1100-1200): the test materials (stone, three opaque lamps of luminance 60, 15 and about 23, a glass lamp
of about 35), LampBrick (air, stone and the four lamp kinds), BorderWorld (bricks around the borders of
8 groups and a 4 x 4 x 4 cluster) and SmallWorld, with the same std::mt19937 draws in the same order.
Brick keys are moved by +16 in y (a whole group: the group borders stay where they are) so that every
key is non-negative, as the world of this build needs."""
from __future__ import annotations

import numpy as np

from ref_fixtures import Mt19937Stream

STONE = 0x8000 | 1
LAMP_BRIGHT = 0x8000 | 5
LAMP_DIM = 0x8000 | 6
LAMP_WARM = 0x8000 | 7
GLASS_LAMP = 9
Y_SHIFT = 16


def lamp_materials() -> np.ndarray:
    """[65536, 8] float32: albedo 0.7, emission per the reference's SetTestMaterials."""
    m = np.zeros((65536, 8), np.float32)
    for code, e in ((STONE, (0, 0, 0)), (LAMP_BRIGHT, (60, 60, 60)), (LAMP_DIM, (15, 15, 15)),
                    (LAMP_WARM, (25, 22, 30)), (GLASS_LAMP, (40, 35, 30))):
        m[code, 0:3] = 0.7
        m[code, 4:7] = e
    return m


def lamp_brick(rng: Mt19937Stream, lamp_percent: int) -> np.ndarray:
    brick = np.zeros(512, np.uint16)
    lamps = (LAMP_BRIGHT, LAMP_DIM, LAMP_WARM, GLASS_LAMP)
    for v in range(512):
        r = rng() % 100
        if r < 50:
            brick[v] = 0
        elif r < 100 - lamp_percent:
            brick[v] = STONE
        else:
            brick[v] = lamps[rng() % 4]
    return brick


def border_world(rng: Mt19937Stream) -> dict:
    world = {}
    for x in range(14, 18):
        for y in range(-2, 2):
            for z in range(30, 34):
                if rng() % 3 != 0:
                    world[(x, y + Y_SHIFT, z)] = lamp_brick(rng, 8)
    for x in range(20, 24):
        for y in range(4, 8):
            for z in range(36, 40):
                world[(x, y + Y_SHIFT, z)] = lamp_brick(rng, 3)
    return world


def small_world(rng: Mt19937Stream) -> dict:
    """x -1..2 moved by +16 (a group) as well."""
    world = {}
    for x in range(-1, 3):
        for y in range(0, 2):
            for z in range(0, 2):
                world[(x + 16, y + Y_SHIFT, z)] = lamp_brick(rng, 6)
    return world


def axis_angle_rows(axis, degrees: float, position) -> np.ndarray:
    """[3, 4] float32 object-to-world rows: the rotation about the unit axis (Rodrigues), then the
    translation."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = np.radians(degrees)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    r = np.eye(3) + np.sin(t) * k + (1 - np.cos(t)) * (k @ k)
    out = np.zeros((3, 4), np.float32)
    out[:, :3] = r
    out[:, 3] = position
    return out
