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
