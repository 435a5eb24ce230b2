"""ecc_normalize -- CPU oracle (test infrastructure only).

Restates ecc_normalize(src, mask, dst, lvl) (core/proc/image_registration/ecc2.cc:385-397):
m = ecc_downscale(src, lvl, BORDER_REPLICATE) (pyrDown lvl times at the default sizes,
ecc2.cc:224-234), ecc_upscale(m, src size) (pyrUp back through the same sizes,
ecc2.cc:237-275), dst = src - m in float, then dst.setTo(0, ~mask): 0 wherever the mask is not
255.  On the FP32 oracle's pyrDown / pyrUp.  Only tests/ use this module.
"""
from __future__ import annotations

import numpy as np

BORDER_REPLICATE = 1


def ecc_normalize(ora, src, mask=None, lvl=1):
    src = np.asarray(src, np.float32)
    h, w = src.shape
    sizes = [(w, h)]
    m = src
    for _ in range(lvl):
        m = ora.pyr_down(m, border=BORDER_REPLICATE)
        sizes.append((m.shape[1], m.shape[0]))
    for j in range(lvl - 1, -1, -1):
        m = ora.pyr_up(m, sizes[j])
    d = (src - m.astype(np.float32)).astype(np.float32)
    if mask is not None:
        d[np.asarray(mask) != 255] = 0
    return d
