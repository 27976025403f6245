"""Measurement setup and voxel grid descriptions (the reference's
DefaultMeasurementSetup, MeasurementSetup.hpp:181-351, and DefaultVoxelGrid,
VoxelGrid.hpp:151-263), with the index maps of MeasurementSetupLinIndex.hpp
and VoxelGridLinIndex.hpp, and the named workloads used by tests and the
benchmark."""
from dataclasses import dataclass

import numpy as np

from . import _capi as capi


@dataclass(frozen=True)
class MeasurementSetup:
    """Two opposing planar heads rotated about y in na steps of da degrees."""
    pos0x: float
    pos1x: float
    na: int
    n0z: int
    n0y: int
    n1z: int
    n1y: int
    da: float
    segx: float
    segy: float
    segz: float

    @property
    def nchannels(self):
        return self.na * self.n0z * self.n0y * self.n1z * self.n1y

    def struct(self):
        return capi.raptr_setup_t(self.pos0x, self.pos1x, self.na, self.n0z, self.n0y,
                                  self.n1z, self.n1y, self.da, self.segx, self.segy, self.segz)

    # MeasurementSetupLinIndex.hpp:78-86 (id1y fastest)
    def lin_id(self, id0z, id0y, id1z, id1y, ida):
        return (((ida * self.n0z + id0z) * self.n0y + id0y) * self.n1z + id1z) * self.n1y + id1y

    def split(self, cnl):
        cnl = np.asarray(cnl, np.int64)
        id1y = cnl % self.n1y
        t = cnl // self.n1y
        id1z = t % self.n1z
        t //= self.n1z
        id0y = t % self.n0y
        t //= self.n0y
        id0z = t % self.n0z
        ida = t // self.n0z
        return ida, id0z, id0y, id1z, id1y


@dataclass(frozen=True)
class VoxelGrid:
    """Axis-aligned grid: origin o, pitch d, counts n (z fastest ids)."""
    o: tuple
    d: tuple
    n: tuple

    @property
    def nvox(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    def struct(self):
        g = capi.raptr_grid_t()
        for k in range(3):
            g.o[k], g.d[k], g.n[k] = float(self.o[k]), float(self.d[k]), int(self.n[k])
        return g

    # VoxelGridLinIndex.hpp:73-80
    def lin_id(self, ix, iy, iz):
        return (ix * self.n[1] + iy) * self.n[2] + iz

    def split(self, v):
        v = np.asarray(v, np.int64)
        iz = v % self.n[2]
        t = v // self.n[2]
        return t // self.n[1], t % self.n[1], iz

    def centres(self):
        ix, iy, iz = self.split(np.arange(self.nvox))
        c = np.empty((self.nvox, 3), np.float64)
        for k, i in enumerate((ix, iy, iz)):
            c[:, k] = self.o[k] + (i + 0.5) * self.d[k]
        return c


def centred_grid(n, pitch):
    half = 0.5 * n * pitch
    return VoxelGrid((-half, -half, -half), (pitch, pitch, pitch), (n, n, n))


def two_head_setup(npix, na, da, pos=230.0, depth=20.0, pitch=4.0):
    return MeasurementSetup(-pos, pos, na, npix, npix, npix, npix, da, depth, pitch, pitch)


# Named workloads (sizes from the benchmark definition of the reference path)
def config_small():
    """8x8-pixel heads (4x4x20 mm) 460 mm apart, 5 angles over 0-180 deg
    (da = 36), 20,480 channels; 32^3 grid at 2.5 mm, centred; 10 rays."""
    return two_head_setup(8, 5, 36.0), centred_grid(32, 2.5), 10


def config_full():
    """13x13-pixel heads, 180 angles at 1 deg (5,140,980 channels); 52^3 grid
    at 2 mm; 30 rays."""
    return two_head_setup(13, 180, 1.0), centred_grid(52, 2.0), 30


def config_tiny():
    """4x4-pixel heads, 3 angles at 60 deg (768 channels); 16^3 grid at 5 mm;
    6 rays (fast parity cases)."""
    return two_head_setup(4, 3, 60.0), centred_grid(16, 5.0), 6
