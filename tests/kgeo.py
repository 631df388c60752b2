"""KGEO geodesic-cache files (apps/compute_kerr_geodesics.cpp:14-22,95-140):
int32 magic 0x4B47454F, version 1, nx, ny, f64 spin; then per ray a 40-byte
header '<5i4xdd' (ray_id, i, j, captured, npoints, alpha, beta) and npoints
PathPoints of 7 f64 {r, theta, x, y, z, dl, redshift}."""
import struct

import numpy as np

MAGIC = 0x4B47454F
HDR = struct.Struct("<iiiid")
RAY = struct.Struct("<5i4xdd")


def write(path, nx, ny, a, rays, magic=MAGIC, version=1):
    """rays: iterable of (ray_id, i, j, captured, alpha, beta, pts[n,7])."""
    with open(path, "wb") as f:
        f.write(HDR.pack(magic, version, nx, ny, a))
        for rid, i, j, cap, al, be, pts in rays:
            pts = np.ascontiguousarray(pts, np.float64)
            f.write(RAY.pack(rid, i, j, cap, pts.shape[0], al, be))
            f.write(pts.tobytes())


def read(path):
    buf = open(path, "rb").read()
    magic, ver, nx, ny, a = HDR.unpack_from(buf, 0)
    pos = HDR.size
    rays = []
    while pos < len(buf):
        rid, i, j, cap, n, al, be = RAY.unpack_from(buf, pos)
        pos += RAY.size
        pts = np.frombuffer(buf, np.float64, n * 7, pos).reshape(n, 7)
        pos += n * 56
        rays.append((rid, i, j, cap, al, be, pts))
    return (magic, ver, nx, ny, a), rays
