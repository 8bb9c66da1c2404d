"""Convert the reference's HDF5 P2 meshes into compressed .npz fixtures.

Run in the development container (reads the reference checkout and the HDF5
C library of this image through ctypes; h5py is not installed):

    python tests/golden/convert_h5_meshes.py

Datasets copied unchanged: /node, /elem, /face, /boundary, /material (the
format read by load_mesh_P2.m:25-57; ids are 0-based).  The element and face
arrays are stored point-major as in the file (n_p x n_e when read in
column-major order).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_MESHES = "/root/reference/slope_stability/meshes"
MESHES = ["SSR_homo_ada_L1.h5"]


def _hdf5():
    for cand in ("/opt/conda/lib/libhdf5.so", "libhdf5.so"):
        try:
            return ctypes.CDLL(cand)
        except OSError:
            continue
    raise OSError("libhdf5 not found")


def read_h5(path):
    h5 = _hdf5()
    h5.H5open()
    hid = ctypes.c_int64
    h5.H5Fopen.restype = hid
    h5.H5Dopen2.restype = hid
    h5.H5Dget_space.restype = hid
    h5.H5Dget_type.restype = hid
    h5.H5Sget_simple_extent_ndims.restype = ctypes.c_int
    h5.H5Tget_class.restype = ctypes.c_int
    h5.H5Tget_size.restype = ctypes.c_size_t
    f = h5.H5Fopen(path.encode(), ctypes.c_uint(0), hid(0))  # H5F_ACC_RDONLY, H5P_DEFAULT
    if f < 0:
        raise OSError(path)
    out = {}
    native_double = hid.in_dll(h5, "H5T_NATIVE_DOUBLE_g").value
    native_llong = hid.in_dll(h5, "H5T_NATIVE_LLONG_g").value
    for name in ("node", "elem", "face", "boundary", "material"):
        d = h5.H5Dopen2(hid(f), ("/" + name).encode(), hid(0))
        if d < 0:
            continue
        sp = h5.H5Dget_space(hid(d))
        nd = h5.H5Sget_simple_extent_ndims(hid(sp))
        dims = (ctypes.c_uint64 * nd)()
        h5.H5Sget_simple_extent_dims(hid(sp), dims, None)
        shape = tuple(int(x) for x in dims)
        t = h5.H5Dget_type(hid(d))
        cls = h5.H5Tget_class(hid(t))  # 0 integer, 1 float
        if cls == 1:
            arr = np.empty(shape, dtype=np.float64)
            mt = native_double
        else:
            arr = np.empty(shape, dtype=np.int64)
            mt = native_llong
        rc = h5.H5Dread(hid(d), hid(mt), hid(0), hid(0), hid(0), arr.ctypes.data_as(ctypes.c_void_p))
        if rc < 0:
            raise OSError(f"H5Dread {name}")
        out[name] = arr
        h5.H5Tclose(hid(t))
        h5.H5Sclose(hid(sp))
        h5.H5Dclose(hid(d))
    h5.H5Fclose(hid(f))
    return out


if __name__ == "__main__":
    if not os.path.isdir(REF_MESHES):
        sys.exit("reference meshes not found")
    for m in MESHES:
        d = read_h5(os.path.join(REF_MESHES, m))
        dst = os.path.join(HERE, "mesh_" + m.replace(".h5", ".npz"))
        np.savez_compressed(dst, **d)
        print(dst, {k: (v.shape, v.dtype) for k, v in d.items()})
