"""Master-frame selection -- CPU oracle (test infrastructure only).

Restates the host rules of select_master_frame's best-of scan
(core/pipeline/c_image_stacking_pipeline_base/c_image_stacking_pipeline_base.cc:312-395) and
the stacking window of create_reference_frame
(core/pipeline/c_image_stacking_pipeline/c_image_stacking_pipeline.cc:1219-1232), and the
per-frame metric: average_bayer_planes for Bayer frames (debayer.cc:277-369), then
compute_local_variance_map's Q (c_local_variance_sharpness_measure.cc:193-249).
Only tests/ use this module.
"""
from __future__ import annotations

from oracle import debayer, lvm


def scan_window(size: int, max_frames_to_scan: int):
    """c_image_stacking_pipeline_base.cc:312-326; the window ends at the middle of the
    sequence (start + max/2), as written there."""
    m = max(3, max_frames_to_scan)
    if size <= m:
        return 0, size
    start = size // 2 - m // 2
    return start, min(size, start + m // 2)


def select(quality, start_pos: int) -> int:
    """c_image_stacking_pipeline_base.cc:369-395: strict improvement over best_metric = 0."""
    best, best_metric = 0, 0.0
    for i, q in enumerate(quality):
        if q > best_metric:
            best, best_metric = i, q
    return best + start_pos


def stack_window(size: int, start_frame_index: int, master_frame_pos: int, max_frames_to_stack: int):
    """c_image_stacking_pipeline.cc:1219-1232."""
    sfi = max(0, start_frame_index)
    if sfi + max_frames_to_stack >= size:
        return sfi, size
    start = max(sfi, master_frame_pos - max_frames_to_stack // 2)
    end = start + max_frames_to_stack
    if end >= size:
        return max(sfi, size - max_frames_to_stack), size
    return start, end


def frame_quality(ora, frame, colorid: int, kradius: int = 1, dscale: int = 1) -> float:
    img = debayer.avgbp(frame) if debayer.is_bayer(colorid) else frame
    return lvm.local_variance_map(ora, img, kradius, dscale)[1]
