"""RabbitCT file formats: the ``.rct`` projection file, the geometry file and
the line-range (clip) cache file.

Layouts (all little-endian, no padding):

* header, 20 bytes: ``uint32 W, H; uint32 N; float HU[2]``
  (ctFileReader_types.h:11-15, read whole at ctFileReader.c:18);
* ``.rct`` record k (k = 0..N-1): 12 float64 (column-major 3x4 matrix, element
  ``a[c*3+r]``) then ``W*H`` float32 (row-major image, ``img[v*W+u]``)
  (ctFileReader.c:51-58, LolaBunny.c:14,64-66);
* geometry file: header then ``N*12`` float64 (ctFileReader.c:36-45);
* clip file: ``N*L*L`` records ``int16 start, stop`` in view-major,
  then ``z*L+y`` order (analyseGeometry.c:159-188,284-292,
  analyseGeometry_types.h:23-26).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

HEADER_FMT = "<3I2f"
HEADER_SIZE = struct.calcsize(HEADER_FMT)  # 20
MATRIX_BYTES = 12 * 8

LINE_RANGE_DTYPE = np.dtype([("start", "<i2"), ("stop", "<i2")])


@dataclass
class RctHeader:
    width: int
    height: int
    n_images: int
    hu: tuple[float, float]

    def pack(self) -> bytes:
        return struct.pack(HEADER_FMT, self.width, self.height, self.n_images,
                           float(self.hu[0]), float(self.hu[1]))

    @classmethod
    def unpack(cls, raw: bytes) -> "RctHeader":
        if len(raw) < HEADER_SIZE:
            raise ValueError("Failed to read file header! Exit.")
        w, h, n, h0, h1 = struct.unpack(HEADER_FMT, raw[:HEADER_SIZE])
        return cls(w, h, n, (h0, h1))

    @property
    def record_bytes(self) -> int:
        return MATRIX_BYTES + 4 * self.width * self.height


def write_rct(path: str | os.PathLike, header: RctHeader, matrices: np.ndarray,
              images) -> None:
    """Write a ``.rct`` file.  ``matrices``: (N,12) float64; ``images``: an
    (N,H,W) float32 array or an iterable yielding N such (H,W) images."""
    matrices = np.ascontiguousarray(matrices, dtype="<f8").reshape(-1, 12)
    if matrices.shape[0] != header.n_images:
        raise ValueError("matrix count does not match the header")
    with open(path, "wb") as f:
        f.write(header.pack())
        for k, img in enumerate(images):
            img = np.ascontiguousarray(img, dtype="<f4")
            if img.shape != (header.height, header.width):
                raise ValueError(f"image {k} has shape {img.shape}")
            f.write(matrices[k].tobytes())
            f.write(img.tobytes())
        if k != header.n_images - 1:
            raise ValueError("image count does not match the header")


def write_geometry(path: str | os.PathLike, header: RctHeader,
                   matrices: np.ndarray) -> None:
    """Write the ``-a`` geometry file: header + N*12 float64."""
    matrices = np.ascontiguousarray(matrices, dtype="<f8").reshape(-1, 12)
    with open(path, "wb") as f:
        f.write(header.pack())
        f.write(matrices.tobytes())


class RctReader:
    """Sequential/random access reader of a ``.rct`` file (memory-mapped).

    ``read_next()`` mirrors ctFileReaderReadImage: it returns the 1-based id of
    the record it read (ctFileReader.c:47-61) and raises on a short file where
    the reference prints "Failed to read projection matrix/image! Exit.".
    """

    def __init__(self, path: str | os.PathLike):
        self.path = os.fspath(path)
        with open(self.path, "rb") as f:
            self.header = RctHeader.unpack(f.read(HEADER_SIZE))
        self._mm = np.memmap(self.path, dtype=np.uint8, mode="r")
        self.current = 0

    def record(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        """(matrix (12,) float64, image (H,W) float32) of record k (0-based)."""
        h = self.header
        off = HEADER_SIZE + k * h.record_bytes
        if off + MATRIX_BYTES > self._mm.size:
            raise IOError("Failed to read projection matrix! Exit.")
        mat = np.frombuffer(self._mm, dtype="<f8", count=12, offset=off)
        if off + h.record_bytes > self._mm.size:
            raise IOError("Failed to read projection image! Exit.")
        img = np.frombuffer(self._mm, dtype="<f4", count=h.width * h.height,
                            offset=off + MATRIX_BYTES).reshape(h.height, h.width)
        return mat, img

    def read_next(self, matrix_out: np.ndarray, image_out: np.ndarray) -> int:
        mat, img = self.record(self.current)
        matrix_out[...] = mat
        image_out.reshape(img.shape)[...] = img
        self.current += 1
        return self.current

    def matrices(self) -> np.ndarray:
        return np.stack([self.record(k)[0] for k in range(self.header.n_images)])


def read_geometry(path: str | os.PathLike) -> tuple[RctHeader, np.ndarray]:
    with open(path, "rb") as f:
        header = RctHeader.unpack(f.read(HEADER_SIZE))
        raw = f.read(header.n_images * MATRIX_BYTES)
    if len(raw) != header.n_images * MATRIX_BYTES:
        raise IOError("Failed to read projection matrix! Exit.")
    return header, np.frombuffer(raw, dtype="<f8").reshape(-1, 12).copy()


def default_clip_filename(L: int, vector_size: int) -> str:
    """``RabbitInput/LineRange<L>-<VS>.rct`` (analyseGeometry.h:9,
    analyseGeometry.c:114-151)."""
    if L not in (128, 256, 512, 1024):
        raise ValueError(f"Unsupported problem size: {L}")
    return f"RabbitInput/LineRange{L}-{vector_size}.rct"


def read_clip_file(path: str | os.PathLike, n_views: int, L: int) -> np.ndarray:
    """(N, L*L, 2) int16 ranges; size-checked like analyseGeometry.c:157-165."""
    size = os.path.getsize(path)
    want = n_views * L * L * 4
    if size != want:
        raise IOError(f"corrupted file {path}: size is {size} (should be {want})")
    return np.fromfile(path, dtype="<i2").reshape(n_views, L * L, 2)


def write_clip_file(path: str | os.PathLike, ranges: np.ndarray) -> None:
    np.ascontiguousarray(ranges, dtype="<i2").tofile(path)
