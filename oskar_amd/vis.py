"""Python handles for the .vis container API (include/oskar_binary.h,
include/oskar_vis.h): binary files, visibility headers and blocks.

Mirrors the reference's oskar/binary/oskar_binary.h and
oskar/vis/oskar_vis_{header,block}.h objects; every call goes through
liboskar_amd.so.
"""
import ctypes as C

import numpy as np

from . import lib
from .mem import _Borrowed, check_status

OSKAR_TAG_GROUP_METADATA = 1
OSKAR_TAG_GROUP_SETTINGS = 3
OSKAR_TAG_GROUP_VIS_HEADER = 11
OSKAR_TAG_GROUP_VIS_BLOCK = 12
OSKAR_ERR_BINARY_TAG_NOT_FOUND = -115
OSKAR_ERR_BINARY_CRC_FAIL = -118


class Binary:
    """An open oskar_Binary file: mode 'w', 'a' or 'r'."""

    def __init__(self, filename, mode, ptr=None):
        if ptr is None:
            st = C.c_int(0)
            ptr = lib.oskar_binary_create(filename.encode(), mode.encode(),
                                          C.byref(st))
            check_status(st, "oskar_binary_create")
        self.ptr = ptr

    def num_tags(self):
        return lib.oskar_binary_num_tags(self.ptr)

    def write(self, data_type, group, tag, index, arr):
        st = C.c_int(0)
        a = np.ascontiguousarray(arr)
        lib.oskar_binary_write(self.ptr, data_type, group, tag, index,
                               a.nbytes, a.ctypes.data_as(C.c_void_p),
                               C.byref(st))
        check_status(st, "oskar_binary_write")

    def read(self, data_type, group, tag, index, dtype):
        """Payload of a chunk as a numpy array of dtype."""
        st = C.c_int(0)
        size = C.c_size_t(0)
        lib.oskar_binary_query(self.ptr, data_type, group, tag, index,
                               C.byref(size), C.byref(st))
        check_status(st, "oskar_binary_query")
        out = np.zeros(size.value, dtype=np.uint8)
        lib.oskar_binary_read(self.ptr, data_type, group, tag, index,
                              size.value, out.ctypes.data_as(C.c_void_p),
                              C.byref(st))
        check_status(st, "oskar_binary_read")
        return out.view(dtype)

    def close(self):
        if self.ptr:
            lib.oskar_binary_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VisHeader:
    """Owning handle to an oskar_VisHeader."""

    def __init__(self, amp_type=0, coord_precision=0, max_times_per_block=0,
                 num_times_total=0, max_channels_per_block=0,
                 num_channels_total=0, num_stations=0, write_autocorr=0,
                 write_crosscorr=1, ptr=None):
        if ptr is None:
            st = C.c_int(0)
            ptr = lib.oskar_vis_header_create(
                amp_type, coord_precision, max_times_per_block,
                num_times_total, max_channels_per_block, num_channels_total,
                num_stations, write_autocorr, write_crosscorr, C.byref(st))
            check_status(st, "oskar_vis_header_create")
        self.ptr = ptr

    @classmethod
    def read(cls, binary):
        st = C.c_int(0)
        ptr = lib.oskar_vis_header_read(binary.ptr, C.byref(st))
        if st.value and ptr:
            lib.oskar_vis_header_free(ptr, C.byref(C.c_int(0)))
        check_status(st, "oskar_vis_header_read")
        return cls(ptr=ptr)

    def write(self, filename):
        """Writes the header chunks; returns the open Binary for blocks."""
        st = C.c_int(0)
        h = lib.oskar_vis_header_write(self.ptr, filename.encode(),
                                       C.byref(st))
        if st.value and h:
            lib.oskar_binary_free(h)
        check_status(st, "oskar_vis_header_write")
        return Binary(None, None, ptr=h)

    def get(self, name):
        return getattr(lib, "oskar_vis_header_" + name)(self.ptr)

    def mem(self, name, *args):
        return _Borrowed(getattr(lib, "oskar_vis_header_" + name)(self.ptr,
                                                                  *args))

    def free(self):
        if self.ptr:
            lib.oskar_vis_header_free(self.ptr, C.byref(C.c_int(0)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class VisBlock:
    """Owning handle to an oskar_VisBlock."""

    def __init__(self, location, amp_type=0, num_times=0, num_channels=0,
                 num_stations=0, create_crosscorr=1, create_autocorr=0,
                 header=None):
        st = C.c_int(0)
        if header is not None:
            self.ptr = lib.oskar_vis_block_create_from_header(
                location, header.ptr, C.byref(st))
        else:
            self.ptr = lib.oskar_vis_block_create(
                location, amp_type, num_times, num_channels, num_stations,
                create_crosscorr, create_autocorr, C.byref(st))
        check_status(st, "oskar_vis_block_create")

    def get(self, name):
        return getattr(lib, "oskar_vis_block_" + name)(self.ptr)

    def cross(self):
        return _Borrowed(lib.oskar_vis_block_cross_correlations(self.ptr))

    def auto(self):
        return _Borrowed(lib.oskar_vis_block_auto_correlations(self.ptr))

    def baseline_uvw(self, dim):
        return _Borrowed(lib.oskar_vis_block_baseline_uvw_metres(self.ptr,
                                                                 dim))

    def station_uvw(self, dim):
        return _Borrowed(lib.oskar_vis_block_station_uvw_metres(self.ptr,
                                                                dim))

    def write(self, binary, block_index):
        st = C.c_int(0)
        lib.oskar_vis_block_write(self.ptr, binary.ptr, block_index,
                                  C.byref(st))
        check_status(st, "oskar_vis_block_write")

    def read(self, header, binary, block_index):
        st = C.c_int(0)
        lib.oskar_vis_block_read(self.ptr, header.ptr, binary.ptr,
                                 block_index, C.byref(st))
        check_status(st, "oskar_vis_block_read")

    def free(self):
        if self.ptr:
            lib.oskar_vis_block_free(self.ptr, C.byref(C.c_int(0)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
