"""Frames in the reference's GSD format through the library's writer (cil_hip_gsd_*).

The reference writes its trajectories and `<name>.Last.gsd` with outputGSDFrame
(OutputWriter.cpp:505-657) through the vendored gsd.c; `GsdWriter` writes the same files.
"""
import ctypes as C

from . import abi


class GsdWriter:
    """A GSD file opened for writing: raw chunks and frames, or whole colony frames."""

    def __init__(self, path, application="CIL", schema="hoomd", schema_version=(1, 1)):
        self.L = abi.lib()
        self.g = C.c_void_p()
        ver = (schema_version[0] << 16) | schema_version[1]
        rc = self.L.cil_hip_gsd_create(str(path).encode(), application.encode(), schema.encode(), ver, C.byref(self.g))
        if rc != 0:
            raise abi.CilError(rc, f"gsd_create {path}")

    def _chk(self, rc, what):
        if rc != 0:
            raise abi.CilError(rc, what)

    def chunk(self, name, type_, data, rows, cols=1):
        """Append one chunk (`data`: a contiguous numpy array of the chunk's type)."""
        self._chk(self.L.cil_hip_gsd_write_chunk(self.g, name.encode(), type_, rows, cols, data.ctypes.data),
                  f"gsd chunk {name}")

    def end_frame(self):
        self._chk(self.L.cil_hip_gsd_end_frame(self.g), "gsd end_frame")

    def frame(self, params, box, col, ts, run=None):
        """One frame of a host colony (outputGSDFrame)."""
        v = col.views()
        self._chk(self.L.cil_hip_gsd_write_frame(self.g, C.byref(params), C.byref(box),
                                                  C.byref(run) if run is not None else None, v, col.nspec, ts),
                  "gsd frame")

    def frame_ctx(self, md, ts, run=None):
        """One frame of a context's device colony."""
        self._chk(self.L.cil_hip_gsd_write_frame_ctx(md.ctx, self.g, C.byref(run) if run is not None else None, ts),
                  "gsd frame")

    def close(self):
        if self.g:
            rc = self.L.cil_hip_gsd_close(self.g)
            self.g = C.c_void_p()
            self._chk(rc, "gsd close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class H5Writer:
    """An HDF5 file in the reference's layout: /params from the run's parameters (or a bare file
    when `params` is None, the reference's `.Last.h5`) and snapshot frames."""

    def __init__(self, path, params=None, box=None, run=None):
        self.L = abi.lib()
        self.f = C.c_void_p()
        rc = self.L.cil_hip_h5_create(str(path).encode(), C.byref(params) if params is not None else None,
                                      C.byref(box) if box is not None else None,
                                      C.byref(run) if run is not None else None, C.byref(self.f))
        if rc != 0:
            raise abi.CilError(rc, f"h5_create {path}")

    def frame(self, params, col, frame_id, ts, run=None, cpu_time=0.0):
        """/snapshots/t_<frame_id> (frame_id < 0: at the root) of a host colony."""
        rc = self.L.cil_hip_h5_write_frame(self.f, C.byref(params), col.views(), col.nspec,
                                           C.byref(run) if run is not None else None, frame_id, ts, cpu_time)
        if rc != 0:
            raise abi.CilError(rc, "h5 frame")

    def frame_ctx(self, md, frame_id, ts, run=None, cpu_time=0.0):
        rc = self.L.cil_hip_h5_write_frame_ctx(md.ctx, self.f, C.byref(run) if run is not None else None, frame_id, ts,
                                               cpu_time)
        if rc != 0:
            raise abi.CilError(rc, "h5 frame")

    def close(self):
        if self.f:
            rc = self.L.cil_hip_h5_close(self.f)
            self.f = C.c_void_p()
            if rc != 0:
                raise abi.CilError(rc, "h5 close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
