"""MVTools motion vectors as search-centre guidance (vnlb's mvfw / mvbw).

The reference reads two frame properties of an MVTools vector clip:
MVTools_MVAnalysisData (21 little-endian 32-bit fields) and MVTools_vectors
(a size header, a validity flag, then the levels of the vector pyramid, of
which the last is the finest, 16 bytes per block: x, y, sad); see
parse_mvtools_analysis / parse_mvtools_vectors (src/flow/flow.hpp:196-318).
A VectorClip holds the finest level of every frame in one dense array that
the GPU walk of MVToolsFlowProvider::center_for (flow.hip) reads.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from .params import VNLBError

_INT_MAX = 2**31 - 1


@dataclass(frozen=True)
class MVToolsAnalysis:
    """The MVTools_MVAnalysisData fields vnlb uses (src/flow/flow.hpp:38-56)."""
    block_size_x: int
    block_size_y: int
    pel: int
    level_count: int
    delta_frame: int
    backwards: bool
    width: int
    height: int
    overlap_x: int
    overlap_y: int
    block_count_x: int
    block_count_y: int
    bits_per_sample: int
    chroma_ratio_x: int
    chroma_ratio_y: int
    padding_x: int
    padding_y: int

    def grid_shape(self) -> dict:
        return {"block_size_x": self.block_size_x, "block_size_y": self.block_size_y,
                "overlap_x": self.overlap_x, "overlap_y": self.overlap_y,
                "block_count_x": self.block_count_x, "block_count_y": self.block_count_y,
                "pel": self.pel}


def parse_mvtools_analysis(data: bytes) -> MVToolsAnalysis:
    """parse_mvtools_analysis (src/flow/flow.hpp:196-258): the same field
    offsets, checks and messages."""
    data = bytes(data)
    if len(data) != 21 * 4:
        raise VNLBError("MVTools_MVAnalysisData must contain 21 32-bit fields")
    u = struct.unpack("<21I", data)

    def positive(index, message):
        if u[index] == 0 or u[index] > _INT_MAX:
            raise VNLBError(message)
        return int(u[index])

    def nonnegative(index, message):
        if u[index] > _INT_MAX:
            raise VNLBError(message)
        return int(u[index])

    analysis = MVToolsAnalysis(
        block_size_x=positive(2, "invalid MVTools horizontal block size"),
        block_size_y=positive(3, "invalid MVTools vertical block size"),
        pel=positive(4, "invalid MVTools pel"),
        level_count=positive(5, "invalid MVTools level count"),
        delta_frame=struct.unpack("<i", data[24:28])[0],
        backwards=u[7] != 0,
        width=positive(10, "invalid MVTools width"),
        height=positive(11, "invalid MVTools height"),
        overlap_x=nonnegative(12, "invalid MVTools horizontal overlap"),
        overlap_y=nonnegative(13, "invalid MVTools vertical overlap"),
        block_count_x=positive(14, "invalid MVTools horizontal block count"),
        block_count_y=positive(15, "invalid MVTools vertical block count"),
        bits_per_sample=positive(16, "invalid MVTools bit depth"),
        chroma_ratio_y=positive(17, "invalid MVTools vertical chroma ratio"),
        chroma_ratio_x=positive(18, "invalid MVTools horizontal chroma ratio"),
        padding_x=nonnegative(19, "invalid MVTools horizontal padding"),
        padding_y=nonnegative(20, "invalid MVTools vertical padding"),
    )
    if analysis.overlap_x >= analysis.block_size_x or analysis.overlap_y >= analysis.block_size_y:
        raise VNLBError("MVTools overlap must be smaller than block size")
    return analysis


def parse_mvtools_vectors(data: bytes, analysis: MVToolsAnalysis):
    """parse_mvtools_vectors (src/flow/flow.hpp:260-318): (valid, vectors
    [block_count_y, block_count_x, 2] int32 (x, y) of the finest level, or
    None when the frame's vectors are flagged invalid)."""
    data = bytes(data)
    if len(data) < 8:
        raise VNLBError("MVTools_vectors is too small")
    declared, valid_flag = struct.unpack_from("<II", data, 0)
    if declared != len(data):
        raise VNLBError("MVTools_vectors size header is invalid")
    if valid_flag != 1:
        return False, None
    position = 8
    finest_begin = finest_end = 0
    while position < len(data):
        if position + 4 > len(data):
            raise VNLBError("truncated MVTools level header")
        (level_size,) = struct.unpack_from("<I", data, position)
        if level_size < 4 or position + level_size > len(data):
            raise VNLBError("invalid MVTools level size")
        finest_begin, finest_end = position + 4, position + level_size
        position = finest_end
    finest = finest_end - finest_begin
    if finest % 16 != 0:
        raise VNLBError("MVTools finest level has invalid vector size")
    if finest // 16 != analysis.block_count_x * analysis.block_count_y:
        raise VNLBError("MVTools finest level vector count does not match analysis data")
    blocks = np.frombuffer(data, dtype="<i4", count=finest // 4, offset=finest_begin).reshape(-1, 4)
    vectors = np.ascontiguousarray(blocks[:, :2], dtype=np.int32)
    return True, vectors.reshape(analysis.block_count_y, analysis.block_count_x, 2)


def scene_change(props: dict, backwards: bool) -> bool:
    """has_scene_change (src/vs/plugin.cpp:476-485) on a source frame's
    properties: _SceneChangeNext for backward vectors, _SceneChangePrev or
    Scenechange for forward ones."""
    if backwards:
        return bool(props.get("_SceneChangeNext", False))
    return bool(props.get("_SceneChangePrev", False)) or bool(props.get("Scenechange", False))


class VectorClip:
    """One MVTools vector clip (mvfw or mvbw) over the frames of a clip:
    vectors [frames, block_count_y, block_count_x, 2] int32 and valid
    [frames] (False where the vectors are flagged invalid or the source frame
    is a scene change; such grids are skipped, as cache_mvtools_grid does,
    src/vs/plugin.cpp:507-523)."""

    def __init__(self, analysis: MVToolsAnalysis, vectors, valid):
        vectors = np.ascontiguousarray(vectors, dtype=np.int32)
        valid = np.ascontiguousarray(valid, dtype=np.uint8)
        expected = (analysis.block_count_y, analysis.block_count_x, 2)
        if vectors.ndim != 4 or vectors.shape[1:] != expected or valid.shape != (vectors.shape[0],):
            raise VNLBError(f"VNLB HIP: vectors must be [frames, {expected[0]}, {expected[1]}, 2] "
                            "with one valid flag per frame")
        self.analysis = analysis
        self.vectors = vectors
        self.valid = valid

    @property
    def frames(self) -> int:
        return int(self.vectors.shape[0])

    @classmethod
    def from_mvtools(cls, analysis_data: bytes, vector_data, source_props=None) -> "VectorClip":
        """From the frame properties of an MVTools vector clip: the
        MVTools_MVAnalysisData blob (read once, from frame 0, as
        read_mvtools_analysis_from_node does) and every frame's
        MVTools_vectors blob; source_props: the source clip's frame
        properties, for the scene-change test."""
        analysis = parse_mvtools_analysis(analysis_data)
        frames = len(vector_data)
        vectors = np.zeros((frames, analysis.block_count_y, analysis.block_count_x, 2), np.int32)
        valid = np.zeros(frames, np.uint8)
        for f, blob in enumerate(vector_data):
            ok, v = parse_mvtools_vectors(blob, analysis)
            if ok and not (source_props is not None and scene_change(source_props[f], analysis.backwards)):
                vectors[f] = v
                valid[f] = 1
        return cls(analysis, vectors, valid)

    def check(self, name: str, frames: int, width: int, height: int, backwards: bool) -> None:
        """validate_mvtools_clip (src/vs/plugin.cpp:452-474)."""
        a = self.analysis
        if self.frames != frames:
            raise VNLBError(f"VNLB HIP: {name} frame count must match clip")
        if a.width != width or a.height != height:
            raise VNLBError(f"VNLB HIP: {name} analysis dimensions must match clip")
        if a.delta_frame != 1:
            raise VNLBError(f"VNLB HIP: {name} must use MVTools delta=1 vectors")
        if a.backwards != backwards:
            raise VNLBError(f"VNLB HIP: {name} must be generated with isb={'True' if backwards else 'False'}")


def analysis_bytes(a: MVToolsAnalysis) -> bytes:
    """The MVTools_MVAnalysisData blob of an analysis (fields 0, 1, 8 and 9,
    which vnlb does not read, are 0)."""
    u = [0] * 21
    u[2], u[3], u[4], u[5] = a.block_size_x, a.block_size_y, a.pel, a.level_count
    u[6] = a.delta_frame & 0xFFFFFFFF
    u[7] = int(a.backwards)
    u[10], u[11], u[12], u[13] = a.width, a.height, a.overlap_x, a.overlap_y
    u[14], u[15], u[16] = a.block_count_x, a.block_count_y, a.bits_per_sample
    u[17], u[18], u[19], u[20] = a.chroma_ratio_y, a.chroma_ratio_x, a.padding_x, a.padding_y
    return struct.pack("<21I", *u)


def vectors_bytes(levels, valid: bool = True) -> bytes:
    """An MVTools_vectors blob: levels is a list of [blocks, 2] int arrays,
    coarsest first (sad fields 0)."""
    body = b""
    for level in levels:
        level = np.asarray(level, dtype=np.int64).reshape(-1, 2)
        records = np.zeros((level.shape[0], 4), dtype="<i4")
        records[:, :2] = level
        payload = records.tobytes()
        body += struct.pack("<I", len(payload) + 4) + payload
    return struct.pack("<II", len(body) + 8, 1 if valid else 0) + body
